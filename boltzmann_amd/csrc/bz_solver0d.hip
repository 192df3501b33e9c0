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
//   * per instance: K profile (sum of the scaled operators, built for the whole
//     batch by k_prepare0d), U rows (WP doubles each, 1/U_jj first).
// Lane/slot mapping: slot s lives in lane s % 64, register level s / 64.  A host
// schedule assigns rows to slots so that no two rows alive at the same step share
// a slot (a row is alive from step first_col - 2W, when its prefetch starts, to its
// pivot step) and uses level 1+ only where more than 64 rows are alive; every
// W-step chunk runs only the levels its rows occupy.
#include <algorithm>
#include "bz_common.h"
#include "../../include/bz.h"

namespace {

using namespace bz;

constexpr int kMaxOps = 16;

struct GridDev {
  int n, ns, npr, bw, S, nop, P, nslots, max_row;
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
  const int* chunk_levels;  // [ceil(n/W)] register levels alive in each W-step chunk
  const unsigned* ent_rc;   // [P] (row << 16) | column of every profile entry
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
  double* vec;              // [B][4][n] scratch (row 0: growth vector of the instance)
  int* iters;               // [B]
  int* status;              // [B]  0 converged, 1 max_iter, 2 line search failed
  double* res_abs;          // [B]
  double* res0;             // [B]
  double atol, rtol, alpha_min, alpha_rho, ls_c;
  int max_iter, mode;       // mode 0: reference Jacobian (d = 0), 1: exact Jacobian
  double tg_scale;          // coefficient of C_Tg is n0 * Tg * tg_scale
  const double* tol_abs;    // [B] per-instance absolute tolerance, or null: atol + rtol * r0
};

__device__ __forceinline__ int orig_index(int i, int ns, int npr) {
  const int q = i / ns, l = i - q * ns;
  return (npr - 1 - q) * ns + l;
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

__device__ __forceinline__ double fast_rcp(double a) {
  double r = __builtin_amdgcn_rcp(a);
  r = fma(r, fma(-a, r, 1.0), r);
  r = fma(r, fma(-a, r, 1.0), r);
  return r;
}

// y = K x over the flat profile: entries are lane-strided (coalesced loads, no
// dependence between iterations) and accumulated into y[row] with LDS atomics.  Groups of
// U entries per lane issue all their loads before the first atomic, so a group waits on
// memory once.  y must be zeroed and synchronised by the caller.
__device__ void flat_matvec(int P, const double* __restrict__ kp, const unsigned* __restrict__ rc,
                            const double* __restrict__ x, double* y, int lane) {
  constexpr int U = 8;
  int e0 = lane;
  for (; e0 + 64 * (U - 1) < P; e0 += 64 * U) {
    unsigned v[U];
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      v[u] = rc[e0 + 64 * u];
      a[u] = kp[e0 + 64 * u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) atomicAdd(&y[v[u] >> 16], a[u] * x[v[u] & 0xffffu]);
  }
  for (int e = e0; e < P; e += 64) {
    const unsigned v = rc[e];
    atomicAdd(&y[v >> 16], kp[e] * x[v & 0xffffu]);
  }
}

// In-place solve M z = b in the permuted order: every l-block (stride ns) is the
// reversed radial mass block with banded Cholesky rows lc[q*4 + d] = L[q, q-d]
// (d = 0 entry stored inverted, out-of-band entries zero) held in LDS; lanes
// 0..ns-1 each own one block and carry the last three values in registers.
__device__ void mass_solve(int ns, int npr, const double* __restrict__ lc, double* __restrict__ b, int lane) {
  if (lane < ns) {
    const int l = lane;
    double y1 = 0.0, y2 = 0.0, y3 = 0.0;
    for (int q = 0; q < npr; ++q) {
      const double* L = lc + q * 4;
      double s = b[q * ns + l];
      s = fma(-L[1], y1, s);
      s = fma(-L[2], y2, s);
      s = fma(-L[3], y3, s);
      const double y = s * L[0];
      b[q * ns + l] = y;
      y3 = y2; y2 = y1; y1 = y;
    }
    double z1 = 0.0, z2 = 0.0, z3 = 0.0;
    for (int q = npr - 1; q >= 0; --q) {
      double s = b[q * ns + l];
      if (q + 1 < npr) s = fma(-lc[(q + 1) * 4 + 1], z1, s);
      if (q + 2 < npr) s = fma(-lc[(q + 2) * 4 + 2], z2, s);
      if (q + 3 < npr) s = fma(-lc[(q + 3) * 4 + 3], z3, s);
      const double z = s * lc[q * 4];
      b[q * ns + l] = z;
      z3 = z2; z2 = z1; z1 = z;
    }
  }
}

__device__ __forceinline__ double dot(const double* __restrict__ a, const double* __restrict__ b, int n, int lane) {
  double s = 0.0;
  for (int i = lane; i < n; i += 64) s = fma(a[i], b[i], s);
  return wave_sum(s);
}

// |(I - uhat uhat^T)(A - mu X)|  (two passes: explicit projection, no |R|^2 - (u.R)^2 cancellation)
__device__ double proj_res_plain(const double* A, const double* X, const double* __restrict__ uh, int n, double mu,
                                 int lane) {
  double su = 0.0;
  for (int i = lane; i < n; i += 64) su = fma(uh[i], A[i] - mu * X[i], su);
  su = wave_sum(su);
  double s2 = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double rp = (A[i] - mu * X[i]) - su * uh[i];
    s2 = fma(rp, rp, s2);
  }
  return sqrt(wave_sum(s2));
}

// Residual norm at x + al*p for the Newton step y = c z1 + d z2, p = y - x, using
// M^-1 K z1 = uhat + mu z1 and M^-1 K z2 = x + mu z2 (the two bordered solves):
//   M^-1 K p = c uhat + d x + mu y - A,   A = M^-1 K x,
//   R(al) = A + al M^-1 K p - mu(al) (x + al p).
__device__ double proj_res_step(const double* A, const double* X, const double* Z1, const double* Z2,
                                const double* __restrict__ uh, int n, double c, double d, double mu, double al,
                                double mua, int lane) {
  double su = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double y = fma(c, Z1[i], d * Z2[i]);
    const double p = y - X[i];
    const double kp = fma(c, uh[i], fma(d, X[i], fma(mu, y, -A[i])));
    const double R = fma(al, kp, A[i]) - mua * fma(al, p, X[i]);
    su = fma(uh[i], R, su);
  }
  su = wave_sum(su);
  double s2 = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double y = fma(c, Z1[i], d * Z2[i]);
    const double p = y - X[i];
    const double kp = fma(c, uh[i], fma(d, X[i], fma(mu, y, -A[i])));
    const double R = fma(al, kp, A[i]) - mua * fma(al, p, X[i]);
    const double rp = R - su * uh[i];
    s2 = fma(rp, rp, s2);
  }
  return sqrt(wave_sum(s2));
}

// y = M x from the M profile (M is l-block diagonal, radial 7-band: columns i + d*ns, |d| <= 3).
__device__ void mass_matvec(int n, int ns, const double* __restrict__ mprof, const int* __restrict__ first_col,
                            const int* __restrict__ poff, const double* __restrict__ x, double* __restrict__ y,
                            int lane) {
  for (int i = lane; i < n; i += 64) {
    const int base = poff[i] - first_col[i];
    double s = 0.0;
#pragma unroll
    for (int d = -3; d <= 3; ++d) {
      const int c = i + d * ns;
      if (c >= 0 && c < n) s = fma(mprof[base + c], x[c], s);
    }
    y[i] = s;
  }
}

// Pointers into LDS, marked as such: a function the compiler keeps out of line would
// otherwise access LDS with generic (flat) instructions, which also wait on global loads.
typedef __attribute__((address_space(3))) double lds_double;
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(3))) short lds_short;
typedef __attribute__((address_space(1))) double glb_double;

// Register state of the factorisation: R levels of row slots per lane.  Each slot also
// holds the row that takes it next (nrow), so that one LDS read of that row's entry gives
// all a slot change needs.  b1/b2 hold the row's two right-hand-side values, eliminated up
// to the current step (read from Y1/Y2 when the row enters).
template <int W, int R>
struct FactorState {
  int row[R], fc[R], off[R], nrow[R];
  unsigned vmask[R];
  double w[R][W], nk[R][W], b1[R], b2[R];
};

// Row table, one int2 per row i: x = first profile column of i (low 16 bits) | (next row
// in i's slot << 16, -1 for none), y = profile offset of i minus that column (entries of
// the row are kprof[y + c]).
__device__ __forceinline__ void fill_rowtab(const GridDev& g, int2* __restrict__ rt, int lane) {
  for (int i = lane; i < g.n; i += 64) {
    const int fc = g.first_col[i];
    int2 v;
    v.x = (int)((unsigned)(fc & 0xffff) | ((unsigned)g.sched_next[i] << 16));
    v.y = g.prof_off[i] - fc;
    rt[i] = v;
  }
}

// Slot state for row `row` (row < 0: an empty slot, first column INT_MAX, never valid).
__device__ __forceinline__ void row_enter(int2 v, int row, int& fc, int& off, int& nrow) {
  fc = (row >= 0) ? (v.x & 0xffff) : 0x7fffffff;
  off = (row >= 0) ? v.y : 0;
  nrow = (row >= 0) ? (v.x >> 16) : -1;
}

// One W-step chunk of the factorisation running levels 0..RA-1 only.  SL: also store the
// multipliers l_rj of L at the profile positions of (r, j) in lout.  TAIL: the chunk may
// end past row n-1 (only the last one); full chunks carry no exit test, so the compiler
// places no spill reloads for an exit path inside the steps (a reload waits on the
// in-order memory counter, i.e. on every column prefetch issued before it).
template <int W, int R, int RA, bool SL, bool TAIL>
__device__ __forceinline__ void factor_chunk(int j0, int n, int bw, FactorState<W, R>& s,
                                             const glb_double* __restrict__ kprof, const lds_int* __restrict__ s_succ,
                                             lds_double* __restrict__ Y1, lds_double* __restrict__ Y2,
                                             lds_double* __restrict__ s_piv, lds_double* __restrict__ s_junk,
                                             glb_double* __restrict__ urows, int WP, int lane,
                                             glb_double* __restrict__ lout) {
#pragma unroll
  for (int t = 0; t < W; ++t) {
    const int j = j0 + t;
    if (TAIL && j >= n) break;
    // --- locate the pivot row j: lane lp, level rr (both wave-uniform)
    int lp = 0, rr = 0;
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const unsigned long long bl = __ballot(s.row[r] == j);
      if (bl) { lp = __ffsll((long long)bl) - 1; rr = r; }
    }
    lp = __builtin_amdgcn_readfirstlane(lp);
    rr = __builtin_amdgcn_readfirstlane(rr);
    const int hit = (lane == lp) ? rr : -1;
    // --- publish U_jj..U_j,j+W-1 and the two rhs values.  Every lane stores (no divergent
    // branch): level rr of the pivot lane into s_piv, the others into their own row of
    // s_junk.
    {
      lds_double* dst = (lane == lp) ? s_piv : s_junk + lane;
#pragma unroll
      for (int r = 0; r < RA; ++r) {
        if (rr == r) {
#pragma unroll
          for (int k = 0; k < W; ++k) dst[k] = s.w[r][(t + k) % W];
          dst[W] = s.b1[r];
          dst[W + 1] = s.b2[r];
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
    // --- U row j (1/U_jj first) and the eliminated rhs; every lane stores entry lane % WP
    {
      const int kk = lane % WP;
      urows[(size_t)j * WP + kk] = (kk == 0) ? rinv : s_piv[kk];
      Y1[j] = pb1;
      Y2[j] = pb2;
    }
    // --- eliminate column j (rows whose first column lies beyond j have a zero there: their
    // window may still hold the previous row of the slot), shift in column j+W (prefetched
    // W steps ago)
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const double m = (j >= s.fc[r]) ? s.w[r][t] * rinv : 0.0;
      if (SL && s.row[r] > j && j >= s.fc[r]) lout[s.off[r] + j] = m;
#pragma unroll
      for (int k = 1; k < W; ++k) s.w[r][(t + k) % W] = fma(-m, pv[k], s.w[r][(t + k) % W]);
      s.b1[r] = fma(-m, pb1, s.b1[r]);
      s.b2[r] = fma(-m, pb2, s.b2[r]);
      s.w[r][t] = ((s.vmask[r] >> t) & 1u) ? s.nk[r][t] : 0.0;
    }
    // --- retire the pivot row; its slot takes the next scheduled row (one LDS read).  Its
    // window is not cleared: the shift-ins of the next W steps overwrite every position
    // (with zero, vmask being clear), and the row's first column is at least 2W steps away.
    if (lane == lp) {
#pragma unroll
      for (int r = 0; r < RA; ++r) {
        if (hit == r) {
          const int nr = s.nrow[r];
          s.row[r] = nr;
          s.vmask[r] = 0u;
          const int q = nr >= 0 ? nr : 0;
          s.b1[r] = (nr >= 0) ? Y1[q] : 0.0;     // entry values (M uhat, M x) of the new row
          s.b2[r] = (nr >= 0) ? Y2[q] : 0.0;
          row_enter(make_int2(s_succ[2 * q], s_succ[2 * q + 1]), nr, s.fc[r], s.off[r], s.nrow[r]);
        }
      }
    }
    // --- branch-free prefetch of column j+2W (invalid columns read entry 0 and are masked);
    // a 32-bit index from the instance's profile base
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int c = j + 2 * W;
      const bool valid = (c >= s.fc[r]) && (c <= min(s.row[r] + bw, n - 1));
      const unsigned idx = valid ? (unsigned)(s.off[r] + c) : 0u;
      s.nk[r][t] = kprof[idx];
      s.vmask[r] = (s.vmask[r] & ~(1u << t)) | ((valid ? 1u : 0u) << t);
    }
  }
}

// Factorise A = K - mu M (the instance profile already holds K - mu M; permuted order,
// no pivoting) and forward-eliminate
// the two right-hand sides b1 = M uhat (rows read from Y1 = M uhat on entry) and
// b2 = M x (read from Y2 on entry).  Y1[j], Y2[j] are overwritten by the eliminated
// values at pivot step j, which is never before the entry of row j.  U rows (1/U_jj
// first) go to urows.  The column prefetch is branch-free so that every step issues
// a fixed number of loads and the memory-counter waits can be counted statically.
template <int W, int R, bool SL = false>
__device__ void factor_forward(const GridDev& g, const double* __restrict__ kprof_,
                               const int2* __restrict__ succ_, const short* __restrict__ lev_,
                               double* __restrict__ y1_, double* __restrict__ y2_, double* __restrict__ piv_,
                               double* __restrict__ junk_, double* __restrict__ urows_, int WP, int lane,
                               double* __restrict__ lout_ = nullptr) {
  const int n = g.n, bw = g.bw;
  const glb_double* kprof = (const glb_double*)kprof_;
  glb_double* urows = (glb_double*)urows_;
  glb_double* lout = (glb_double*)lout_;
  const lds_int* s_succ = (const lds_int*)succ_;          // the row table as int pairs
  const lds_short* s_lev = (const lds_short*)lev_;
  lds_double* Y1 = (lds_double*)y1_;
  lds_double* Y2 = (lds_double*)y2_;
  lds_double* s_piv = (lds_double*)piv_;
  lds_double* s_junk = (lds_double*)junk_;
  FactorState<W, R> s;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int rw = g.slot_first[r * 64 + lane];
    s.row[r] = rw;
    const int lc = (rw >= 0) ? min(rw + bw, n - 1) : -1;
    const int q = rw >= 0 ? rw : 0;
    row_enter(make_int2(s_succ[2 * q], s_succ[2 * q + 1]), rw, s.fc[r], s.off[r], s.nrow[r]);
    s.b1[r] = (rw >= 0) ? Y1[q] : 0.0;
    s.b2[r] = (rw >= 0) ? Y2[q] : 0.0;
    s.vmask[r] = 0u;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const int c0 = k, c1 = W + k;
      double v0 = 0.0, a1 = 0.0;
      if (c0 >= s.fc[r] && c0 <= lc) v0 = kprof[s.off[r] + c0];
      if (c1 >= s.fc[r] && c1 <= lc) {
        a1 = kprof[s.off[r] + c1];
        s.vmask[r] |= 1u << k;
      }
      s.w[r][k] = v0; s.nk[r][k] = a1;
    }
  }
  // Wait for the initial window loads here, once: with them still in flight at the loop
  // entry, the compiler's wait analysis would put a full vector-memory wait in every chunk.
  // (gfx9 s_waitcnt fields: vmcnt 0, expcnt 7, lgkmcnt 15.)
  __builtin_amdgcn_s_waitcnt(0x0F70);
  // Y1/Y2 double as the entry values of the rhs (M uhat, M x): a row entering its
  // slot during step j has index > j, and Y*[i] is only overwritten at pivot step i.
  const int nfull = n / W;
  for (int ch = 0; ch < nfull; ++ch)
    factor_chunk<W, R, R, SL, false>(ch * W, n, bw, s, kprof, s_succ, Y1, Y2, s_piv, s_junk, urows, WP, lane,
                                       lout);
  // last partial chunk: all levels (running a level with no rows is harmless)
  if (nfull * W < n)
    factor_chunk<W, R, R, SL, true>(nfull * W, n, bw, s, kprof, s_succ, Y1, Y2, s_piv, s_junk, urows, WP, lane,
                                     lout);
}

// Back substitution U z = y for the two right-hand sides held in Y1/Y2 (in place).
// U rows are staged through LDS in chunks of 2W rows: all 64 lanes load the next chunk
// (contiguous, coalesced) into registers while lanes 0 and 1 (one right-hand side each)
// sweep the current chunk from LDS; the window z[] holds z_{j+1..j+W-1} at static
// positions j mod W (chunk bases are multiples of W).
template <int W>
__device__ void back_substitute(int n, const double* __restrict__ urows, int WP, double* __restrict__ Y1,
                                double* __restrict__ Y2, double* __restrict__ stg_, int lane) {
  constexpr int CH = 2 * W;
  constexpr int PER = (CH * ((W + 1) & ~1) + 63) / 64;   // doubles per lane per chunk
  const int csize = CH * WP;
  // Y1, Y2 and stg are in LDS: say so, so that the accesses are LDS instructions (counted
  // apart from the global loads of the next chunk) rather than generic ones
  typedef __attribute__((address_space(3))) double lds_double;
  lds_double* Y = (lds_double*)(lane == 0 ? Y1 : Y2);
  lds_double* stg = (lds_double*)stg_;
  double z[W];
#pragma unroll
  for (int k = 0; k < W; ++k) z[k] = 0.0;
  const int nch = (n + CH - 1) / CH;
  double buf[PER];
  {
    const int base = (nch - 1) * CH * WP;
    const int lim = n * WP;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int o = lane + 64 * e;
      const int gi = base + o;
      buf[e] = urows[(o < csize && gi < lim) ? gi : 0];
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int o = lane + 64 * e;
      if (o < csize) stg[o] = buf[e];
    }
  }
  __syncthreads();
  for (int c = nch - 1; c >= 0; --c) {
    if (c > 0) {
      const int base = (c - 1) * CH * WP;
#pragma unroll
      for (int e = 0; e < PER; ++e) {
        const int o = lane + 64 * e;
        buf[e] = urows[base + (o < csize ? o : 0)];
      }
    }
    if (lane < 2) {
      const int j0 = c * CH;
#pragma unroll
      for (int tt = CH - 1; tt >= 0; --tt) {
        const int j = j0 + tt;
        if (j < n) {
          const lds_double* ur = stg + tt * WP;
          double sum = Y[j];
#pragma unroll
          for (int k = W - 1; k >= 1; --k) sum = fma(-ur[k], z[(tt + k) % W], sum);
          const double zj = sum * ur[0];
          z[tt % W] = zj;
          Y[j] = zj;
        }
      }
    }
    __syncthreads();
    if (c > 0) {
#pragma unroll
      for (int e = 0; e < PER; ++e) {
        const int o = lane + 64 * e;
        if (o < csize) stg[o] = buf[e];
      }
    }
    __syncthreads();
  }
}

// Per-instance K profile (sum of the scaled operators) and growth vector, all instances at once.
__global__ void k_prepare0d(GridDev g, BatchDev b) {
  const long total = (long)b.B * g.P;
  const int nk = g.nop - 1;
  const int S = g.S;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int inst = (int)(idx / g.P);
    const int e = (int)(idx - (long)inst * g.P);
    const double n0 = b.n0[inst];
    double v = 0.0;
    for (int m = 0; m < nk; ++m) {
      double cm;
      if (m < S) cm = n0 * b.ns_by_n0[(size_t)m * b.B + inst];
      else if (m == S) cm = n0 * b.Tg[inst] * b.tg_scale;
      else cm = b.E[inst];
      v = fma(cm, g.opprof[(size_t)m * g.P + e], v);
    }
    b.kprof[idx] = v;
  }
  const long totw = (long)b.B * g.n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < totw; idx += (long)gridDim.x * blockDim.x) {
    const int inst = (int)(idx / g.n);
    const int i = (int)(idx - (long)inst * g.n);
    const double n0 = b.n0[inst];
    double v = 0.0;
    for (int m = 0; m < nk; ++m) {
      double cm;
      if (m < S) cm = n0 * b.ns_by_n0[(size_t)m * b.B + inst];
      else if (m == S) cm = n0 * b.Tg[inst] * b.tg_scale;
      else cm = b.E[inst];
      v = fma(cm, g.wsp[(size_t)m * g.n + i], v);
    }
    b.vec[(size_t)inst * 4 * g.n + i] = v;
  }
}

// LDS layout (doubles first): X, Z1, Z2 (+ pad so Z1..Z2 can hold 4*npr mass Cholesky
// coefficients), AV, STG[2W*WP], PIV[PIVN]; then SC (int2, row table), PO (int),
// LEV (short).
__host__ __device__ __forceinline__ int zpad(int n, int npr) { return (4 * npr > 2 * n) ? (4 * npr - 2 * n) : 0; }

// Two waves per SIMD (register budget 256 incl. spills) measured faster than one
// wave holding every value in registers.
template <int W, int R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W >= 17 ? 1 : 2))) void k_newton0d(GridDev g, BatchDev b) {
  extern __shared__ double lds[];
  const int n = g.n, lane = threadIdx.x, inst = blockIdx.x;
  constexpr int WP = (W + 1) & ~1;
  constexpr int PIVN = (W + 3) & ~1;
  const int nchunk = (n + W - 1) / W;
  double* X = lds;
  double* Z1 = X + n;
  double* Z2 = Z1 + n;
  double* AV = Z2 + n + zpad(n, g.npr);                    // M^-1 K x
  double* STG = AV + n;                                    // back-substitution stage
  double* PIV = STG + 2 * W * WP;                          // published pivot row
  int2* SC = reinterpret_cast<int2*>(PIV + PIVN);
  int* PO = reinterpret_cast<int*>(SC + n);
  short* LEV = reinterpret_cast<short*>(PO + (n + 1));
  double* LCH = Z1;                                        // staged only when Z1/Z2 are free

  double* kprof = b.kprof + (size_t)inst * g.P;           // holds K - mu_app M
  const double* __restrict__ mprof = g.opprof + (size_t)(g.nop - 1) * g.P;
  double mu_app = 0.0;
  const double* __restrict__ UH = g.uhat;
  const double* __restrict__ WV = b.vec + (size_t)inst * 4 * n;
  double* urows = b.urows + (size_t)inst * n * WP;

  // ---------------- tables into LDS, initial iterate (original (n,B) layout -> permuted)
  fill_rowtab(g, SC, lane);
  for (int i = lane; i < n; i += 64) X[i] = b.x[(size_t)orig_index(i, g.ns, g.npr) * b.B + inst];
  for (int i = lane; i <= n; i += 64) PO[i] = g.prof_off[i];
  for (int i = lane; i < nchunk; i += 64) LEV[i] = (short)g.chunk_levels[i];
  __syncthreads();
  {
    double sacc = 0.0;
    for (int i = lane; i < n; i += 64) sacc = fma(g.u[i], X[i], sacc);
    const double ux = wave_sum(sacc);
    for (int i = lane; i < n; i += 64) X[i] = X[i] / ux;
  }

  // residual A = M^-1 K x from scratch (uses Z1/Z2 as the coefficient stage)
  auto true_residual = [&]() {
    for (int e = lane; e < 4 * g.npr; e += 64) {
      const int q = e >> 2, dd = e & 3;
      const double v = g.mchol[(size_t)(q * g.ns) * 4 + dd];
      LCH[e] = (dd == 0) ? 1.0 / v : ((q - dd >= 0) ? v : 0.0);
    }
    for (int i = lane; i < n; i += 64) AV[i] = 0.0;
    __syncthreads();
    flat_matvec(g.P, kprof, g.ent_rc, X, AV, lane);
    __syncthreads();
    mass_solve(g.ns, g.npr, LCH, AV, lane);
    __syncthreads();
    for (int i = lane; i < n; i += 64) AV[i] = fma(mu_app, X[i], AV[i]);   // M^-1 (K - mu_app M + mu_app M) x
    __syncthreads();
  };

  true_residual();
  double mu = dot(WV, X, n, lane);
  const double r0 = proj_res_plain(AV, X, UH, n, mu, lane);
  double rcur = r0;
  double abs_err = r0;
  const double tol = b.tol_abs ? b.tol_abs[inst] : b.atol + b.rtol * r0;
  int it = 0, status = 1, verify = 0;

  for (;;) {
    bool done = false;
    if (b.mode == 0) { if (!(abs_err > tol)) { status = 0; done = true; } }
    else if (!(rcur > tol)) { status = 0; done = true; }
    if (!done && it >= b.max_iter) { status = 1; done = true; }
    if (done) {
      // recompute the residual from scratch; iterate on if the incremental value drifted
      true_residual();
      mu = dot(WV, X, n, lane);
      const double rtrue = proj_res_plain(AV, X, UH, n, mu, lane);
      rcur = rtrue;
      if (status == 0 && rtrue > tol && it < b.max_iter && verify < 2) {
        ++verify;
        abs_err = rtrue;
        status = 1;
      } else {
        if (status == 0 && rtrue > tol) status = 1;
        break;
      }
    }
    if (b.mode == 0) abs_err = rcur;
    mu = dot(WV, X, n, lane);
    {
      // one pass over the rows: shift the profile from K - mu_app M to K - mu M on the band
      // of M (row i: columns i + d*ns, |d| <= 3) and form the right-hand sides M uhat, M x.
      // The row table gives the row's profile offset minus its first column; all band values
      // of a row are loaded before any store, so they are in flight together.
      const double dmu = mu - mu_app;
      for (int i = lane; i < n; i += 64) {
        const int base = SC[i].y;
        double kv[7], mv[7];
#pragma unroll
        for (int d = 0; d < 7; ++d) {
          const int c = i + (d - 3) * g.ns;
          const bool ok = (c >= 0) && (c < n);
          const int cc = ok ? c : i;                 // the diagonal: always in the row
          mv[d] = ok ? mprof[base + cc] : 0.0;
          kv[d] = kprof[base + cc];
        }
        double sacc = 0.0;
#pragma unroll
        for (int d = 0; d < 7; ++d) {
          const int c = i + (d - 3) * g.ns;
          if ((c >= 0) && (c < n)) {
            sacc = fma(mv[d], X[c], sacc);
            kprof[base + c] = fma(-dmu, mv[d], kv[d]);
          }
        }
        Z1[i] = g.muhat[i];
        Z2[i] = sacc;
      }
      mu_app = mu;
    }
    __syncthreads();
    factor_forward<W, R>(g, kprof, SC, LEV, Z1, Z2, PIV, STG, urows, WP, lane);
    __syncthreads();
    back_substitute<W>(n, urows, WP, Z1, Z2, STG, lane);
    __syncthreads();
    // --- 2x2 system for (c, d)
    const double s11 = dot(UH, Z1, n, lane);
    const double s12 = dot(UH, Z2, n, lane);
    const double ux = dot(UH, X, n, lane);
    const double wz1 = dot(WV, Z1, n, lane);
    const double wz2 = dot(WV, Z2, n, lane);
    double c, d;
    if (b.mode == 0) {
      c = ux / s11;
      d = 0.0;
    } else {
      const double s22 = wz2 - 1.0;
      const double det = s11 * s22 - s12 * wz1;
      c = (ux * s22 - s12 * mu) / det;
      d = (s11 * mu - wz1 * ux) / det;
    }
    const double mup = c * wz1 + d * wz2 - mu;         // w.p
    // --- Armijo backtracking on 0.5 |r|^2 (bte_0d3v_batched.py:936-952)
    const double ae2 = rcur * rcur;
    double alpha = 1.0;
    bool diverged = false;
    double rnew = proj_res_step(AV, X, Z1, Z2, UH, n, c, d, mu, alpha, mu + alpha * mup, lane);
    while (0.5 * rnew * rnew > 0.5 * ae2 - b.ls_c * alpha * ae2) {
      alpha *= b.alpha_rho;
      if (alpha < b.alpha_min) { diverged = true; break; }
      rnew = proj_res_step(AV, X, Z1, Z2, UH, n, c, d, mu, alpha, mu + alpha * mup, lane);
    }
    if (diverged) {
      status = 2;
      break;
    }
    for (int i = lane; i < n; i += 64) {
      const double y = fma(c, Z1[i], d * Z2[i]);
      const double p = y - X[i];
      const double kp = fma(c, UH[i], fma(d, X[i], fma(mu, y, -AV[i])));
      X[i] = fma(alpha, p, X[i]);
      AV[i] = fma(alpha, kp, AV[i]);
    }
    __syncthreads();
    rcur = rnew;
    if (b.mode == 0 && (it % 10) == 0) abs_err = rcur;
    ++it;
  }

  // ---------------- write back (original order)
  for (int i = lane; i < n; i += 64) b.x[(size_t)orig_index(i, g.ns, g.npr) * b.B + inst] = X[i];
  if (lane == 0) {
    b.iters[inst] = it;
    b.status[inst] = status;
    b.res_abs[inst] = rcur;
    b.res0[inst] = r0;
  }
}

template <int W>
size_t newton_lds_bytes(int n, int npr) {
  constexpr int WP = (W + 1) & ~1;
  constexpr int PIVN = (W + 3) & ~1;
  const int nchunk = (n + W - 1) / W;
  return (size_t)(4 * n + zpad(n, npr) + 2 * W * WP + PIVN) * sizeof(double) + (size_t)n * sizeof(int2) +
         (size_t)(n + 1) * sizeof(int) + (size_t)nchunk * sizeof(short);
}

struct EEDev {
  const double* ccA;    // [npr][7][7][ns][ns][ns]
  const double* ccB;
  const double* h;      // (B, n) original order
  const double* g;
  const double* negam;  // [B]
  const double* h_lo;   // (B, n) or null
  const double* g_lo;
};

// K profile += ne gamma_a [cc_a(., h) + cc_b(., g)]: the e-e operator of every instance at
// its current potentials, entry (row (p,q), column (k,l)) = sum_{r,s} cc[(p,q),(k,l),(r,s)] h[(r,s)]
// (the M-space form of cc = M^-1 (cc_a hl_op + cc_b gl_op), bte_0d3v_batched.py:462-468).
__global__ void k_ee_profile_add(GridDev g, BatchDev b, EEDev e) {
  const long total = (long)b.B * g.P;
  const int ns = g.ns, n = g.n, npr = g.npr;
  const int nc = ns * ns * ns;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int inst = (int)(idx / g.P);
    const int ent = (int)(idx - (long)inst * g.P);
    const unsigned rc = g.ent_rc[ent];
    const int oi = orig_index((int)(rc >> 16), ns, npr), oj = orig_index((int)(rc & 0xffffu), ns, npr);
    const int p = oi / ns, q = oi - p * ns, k = oj / ns, l = oj - k * ns;
    const int dk = k - p;
    if (dk < -3 || dk > 3) continue;
    const double* hv = e.h + (size_t)inst * n;
    const double* gv = e.g + (size_t)inst * n;
    const double* hl = e.h_lo ? e.h_lo + (size_t)inst * n : nullptr;
    const double* gl = e.g_lo ? e.g_lo + (size_t)inst * n : nullptr;
    // Compensated sum: near v_max the products cancel to ~1e-6 of their size, so a plain
    // FP64 sum carries order-dependent errors that bias the e-e fixed point.
    double s = 0.0, c = 0.0;
    for (int dr = -3; dr <= 3; ++dr) {
      const int r = p + dr;
      if (r < 0 || r >= npr) continue;
      const size_t base = (((size_t)p * 7 + (dk + 3)) * 7 + (dr + 3)) * nc + (size_t)(q * ns + l) * ns;
      for (int ss = 0; ss < ns; ++ss) {
        two_sum_acc(e.ccA[base + ss], hv[r * ns + ss], s, c);
        two_sum_acc(e.ccB[base + ss], gv[r * ns + ss], s, c);
        if (hl) c = fma(e.ccA[base + ss], hl[r * ns + ss], fma(e.ccB[base + ss], gl[r * ns + ss], c));
      }
    }
    b.kprof[idx] = fma(e.negam[inst], s + c, b.kprof[idx]);
  }
}

// ---------------------------------------------------------------- e-e Newton-Krylov pieces
// The Newton system with e-e, J d = -F with
//   F(x)  = P (M^-1 K_x x - mu x),      K_x = K + ne gamma_a [cc_a(., Hx) + cc_b(., Gx)]
//   J v   = P (M^-1 (K_x v + T2 v) - mu v - x (w.v)),   T2 v = ne gamma_a [cc_a(x, Hv) + cc_b(x, Gv)]
// (P = I - uhat uhat^T; the reference's Jacobian, bte_0d3v_batched.py:734-743, plus the
// growth derivative) is solved by GMRES on the host with the banded part as right
// preconditioner; these kernels are its building blocks.  The instance profile holds
// K_x - mu M (mu = w.x of the Newton iterate), one wavefront per instance.

// mu[b] = w.x_b and the profile shift K_x -> K_x - mu M
__global__ __launch_bounds__(64) void k_ee_mu_shift(GridDev g, BatchDev b, double* __restrict__ mu_out) {
  const int n = g.n, lane = threadIdx.x, inst = blockIdx.x;
  const double* WV = b.vec + (size_t)inst * 4 * n;
  double s = 0.0;
  for (int i = lane; i < n; i += 64) s = fma(WV[i], b.x[(size_t)orig_index(i, g.ns, g.npr) * b.B + inst], s);
  const double mu = wave_sum(s);
  if (lane == 0) mu_out[inst] = mu;
  double* kprof = b.kprof + (size_t)inst * g.P;
  const double* __restrict__ mprof = g.opprof + (size_t)(g.nop - 1) * g.P;
  for (int k = lane; k < 7 * n; k += 64) {
    const int i = k / 7, dd = k - 7 * (k / 7) - 3;
    const int c = i + dd * g.ns;
    if (c >= 0 && c < n) {
      const int e = g.prof_off[i] + c - g.first_col[i];
      kprof[e] = fma(-mu, mprof[e], kprof[e]);
    }
  }
}

// Stage the mass Cholesky rows (LCH, 4*npr doubles) used by mass_solve.
__device__ void stage_mass(const GridDev& g, double* LCH, int lane) {
  for (int e = lane; e < 4 * g.npr; e += 64) {
    const int q = e >> 2, dd = e & 3;
    const double v = g.mchol[(size_t)(q * g.ns) * 4 + dd];
    LCH[e] = (dd == 0) ? 1.0 / v : ((q - dd >= 0) ? v : 0.0);
  }
}

// mode 0: out = F(x), nrm[b] = |F|;  mode 1: out = J v (v, t2 given, (B, n) original order).
// Instances with active[b] == 0 (when active is given) are skipped.
__global__ __launch_bounds__(64) void k_ee_apply(GridDev g, BatchDev b, int mode, const int* __restrict__ active,
                                                 const double* __restrict__ v, const double* __restrict__ t2,
                                                 double* __restrict__ out, double* __restrict__ nrm) {
  extern __shared__ double lds[];
  const int n = g.n, lane = threadIdx.x, inst = blockIdx.x;
  if (active && !active[inst]) return;
  double* X = lds;
  double* V = X + n;
  double* AV = V + n;
  double* LCH = AV + n;
  const double* kprof = b.kprof + (size_t)inst * g.P;
  const double* WV = b.vec + (size_t)inst * 4 * n;
  const double* UH = g.uhat;
  for (int i = lane; i < n; i += 64) {
    const int o = orig_index(i, g.ns, g.npr);
    X[i] = b.x[(size_t)o * b.B + inst];
    V[i] = mode ? v[(size_t)inst * n + o] : X[i];
    AV[i] = mode ? t2[(size_t)inst * n + o] : 0.0;
  }
  stage_mass(g, LCH, lane);
  __syncthreads();
  flat_matvec(g.P, kprof, g.ent_rc, V, AV, lane);                 // (K_x - mu M) v (+ T2 v)
  __syncthreads();
  mass_solve(g.ns, g.npr, LCH, AV, lane);                         // M^-1 (...)
  __syncthreads();
  double wv = 0.0;
  if (mode) wv = dot(WV, V, n, lane);
  double su = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double r = mode ? fma(-wv, X[i], AV[i]) : AV[i];
    AV[i] = r;
    su = fma(UH[i], r, su);
  }
  su = wave_sum(su);
  double s2 = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double r = AV[i] - su * UH[i];
    out[(size_t)inst * n + orig_index(i, g.ns, g.npr)] = r;
    s2 = fma(r, r, s2);
  }
  s2 = wave_sum(s2);
  if (lane == 0 && nrm) nrm[inst] = sqrt(s2);
}

// The preconditioner with the factorisation kept: once per Newton step, L (multipliers, in
// the profile positions of the instance, lprof) and U (urows) of K_x - mu M, and
// z1 = (K_x - mu M)^-1 M uhat with s1 = uhat.z1 (permuted order); then every application
// is a mass product, one profile forward solve and one back substitution.
template <int W, int R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W >= 17 ? 1 : 2))) void k_ee_factor(
    GridDev g, BatchDev b, const int* __restrict__ active, double* __restrict__ lprof, double* __restrict__ z1,
    double* __restrict__ s1) {
  extern __shared__ double lds[];
  const int n = g.n, lane = threadIdx.x, inst = blockIdx.x;
  if (active && !active[inst]) return;
  constexpr int WP = (W + 1) & ~1;
  constexpr int PIVN = (W + 3) & ~1;
  const int nchunk = (n + W - 1) / W;
  double* Z1 = lds + n;
  double* Z2 = Z1 + n;
  double* AV = Z2 + n + zpad(n, g.npr);
  double* STG = AV + n;
  double* PIV = STG + 2 * W * WP;
  int2* SC = reinterpret_cast<int2*>(PIV + PIVN);
  short* LEV = reinterpret_cast<short*>(SC + n);
  const double* kprof = b.kprof + (size_t)inst * g.P;
  double* urows = b.urows + (size_t)inst * n * WP;
  fill_rowtab(g, SC, lane);
  for (int i = lane; i < n; i += 64) {
    Z1[i] = g.muhat[i];
    Z2[i] = 0.0;
  }
  for (int i = lane; i < nchunk; i += 64) LEV[i] = (short)g.chunk_levels[i];
  __syncthreads();
  factor_forward<W, R, true>(g, kprof, SC, LEV, Z1, Z2, PIV, STG, urows, WP, lane, lprof + (size_t)inst * g.P);
  __syncthreads();
  back_substitute<W>(n, urows, WP, Z1, Z2, STG, lane);
  __syncthreads();
  const double su = dot(g.uhat, Z1, n, lane);
  for (int i = lane; i < n; i += 64) z1[(size_t)inst * n + i] = Z1[i];
  if (lane == 0) s1[inst] = su;
}

// Forward solve with the stored unit lower factor, row by row; the profile rows are staged
// through LDS in chunks of whole rows (at most LBUF entries) with the next chunk's loads in
// flight (registers) while the current chunk is swept, so the sweep waits on LDS only.
constexpr int LBUF = 1024;

template <int W>
__global__ __launch_bounds__(64) void k_ee_precond_apply(GridDev g, BatchDev b, const int* __restrict__ active,
                                                         const double* __restrict__ lprof,
                                                         const double* __restrict__ z1, const double* __restrict__ s1,
                                                         const double* __restrict__ v, double* __restrict__ out,
                                                         double* __restrict__ outT) {
  extern __shared__ double lds[];
  const int n = g.n, lane = threadIdx.x, inst = blockIdx.x;
  if (active && !active[inst]) return;
  constexpr int WP = (W + 1) & ~1;
  constexpr int PER = LBUF / 64;
  double* VV = lds;
  double* Y1 = VV + n;
  double* Y2 = Y1 + n;
  double* STG = Y2 + n;
  double* SB = STG + 2 * W * WP;                 // two stages of LBUF
  int* PO = reinterpret_cast<int*>(SB + 2 * LBUF);
  short* FC = reinterpret_cast<short*>(PO + (n + 1));
  const double* __restrict__ mprof = g.opprof + (size_t)(g.nop - 1) * g.P;
  const double* __restrict__ L = lprof + (size_t)inst * g.P;
  const double* urows = b.urows + (size_t)inst * n * WP;
  for (int i = lane; i < n; i += 64) {
    VV[i] = v[(size_t)inst * n + orig_index(i, g.ns, g.npr)];
    Y1[i] = 0.0;
    FC[i] = (short)g.first_col[i];
  }
  for (int i = lane; i <= n; i += 64) PO[i] = g.prof_off[i];
  __syncthreads();
  mass_matvec(n, g.ns, mprof, g.first_col, g.prof_off, VV, Y2, lane);
  // rows [r0, r_end(r0)) fit one stage (every row has at most LBUF entries: checked at launch)
  auto r_end = [&](int r0) {
    int r1 = r0 + 1;
    while (r1 < n && PO[r1 + 1] - PO[r0] <= LBUF) ++r1;
    return r1;
  };
  double reg[PER];
  int r0 = 0, r1 = r_end(0), st = 0;
  {
    const int e0 = PO[0], cnt = PO[r1] - e0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int o = lane + 64 * k;
      if (o < cnt) SB[o] = L[e0 + o];
    }
  }
  __syncthreads();
  while (r0 < n) {
    const int r2 = (r1 < n) ? r_end(r1) : n;
    const int e1 = PO[r1], cnt1 = PO[r2] - e1;    // cnt1 = 0 after the last stage
    if (cnt1 > 0) {
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int o = lane + 64 * k;
        reg[k] = L[e1 + (o < cnt1 ? o : 0)];
      }
    }
    const double* S = SB + st * LBUF;
    const int base = PO[r0];
    if (PO[r1] - base > 128 * (r1 - r0)) {
      // long rows (more than 128 entries on average, few rows per stage): each row's sum over
      // its columns left of the diagonal is split across the 64 lanes and reduced
      for (int i = r0; i < r1; ++i) {
        const int fc = FC[i], off = PO[i] - base - fc;
        double acc = 0.0;
        for (int j = fc + lane; j < i; j += 64) acc = fma(-S[off + j], Y2[j], acc);
        acc = wave_sum(acc);
        if (lane == 0) Y2[i] += acc;
        __syncthreads();
      }
    } else
    // blocks of 64 rows, lane t owning row blk0 + t: first the columns left of the block
    // (values final), then the unit lower triangle of the block, one column per step with
    // the finished value broadcast from its lane
    for (int blk0 = r0; blk0 < r1; blk0 += 64) {
      const int nb = min(64, r1 - blk0);
      const int i = blk0 + lane;
      const bool act = lane < nb;
      const int fc = act ? (int)FC[i] : 0;
      const int off = act ? PO[i] - base - fc : 0;
      double acc = act ? Y2[i] : 0.0;
      if (act)
        for (int j = fc; j < blk0; ++j) acc = fma(-S[off + j], Y2[j], acc);
      for (int t = 0; t < nb; ++t) {
        const int j = blk0 + t;
        const int lo = __builtin_amdgcn_readlane(__double2loint(acc), t);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(acc), t);
        const double yj = __hiloint2double(hi, lo);
        if (act && i > j && j >= fc) acc = fma(-S[off + j], yj, acc);
      }
      if (act) Y2[i] = acc;
      __syncthreads();
    }
    double* D = SB + (st ^ 1) * LBUF;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int o = lane + 64 * k;
      if (o < cnt1) D[o] = reg[k];
    }
    __syncthreads();
    st ^= 1;
    r0 = r1;
    r1 = r2;
  }
  back_substitute<W>(n, urows, WP, Y1, Y2, STG, lane);
  __syncthreads();
  const double c = -dot(g.uhat, Y2, n, lane) / s1[inst];
  const double* Z1 = z1 + (size_t)inst * n;
  for (int i = lane; i < n; i += 64) {
    const int o = orig_index(i, g.ns, g.npr);
    const double z = fma(c, Z1[i], Y2[i]);
    out[(size_t)inst * n + o] = z;
    if (outT) outT[(size_t)o * b.B + inst] = z;
  }
}

template <int W, int R>
int launch_ee_factor(const GridDev& g, const BatchDev& b, const int* active, double* lprof, double* z1, double* s1,
                     hipStream_t stream) {
  const size_t lds = newton_lds_bytes<W>(g.n, g.npr);
  if (lds > 64 * 1024) return BZ_EINVAL;
  hipLaunchKernelGGL((k_ee_factor<W, R>), dim3(b.B), dim3(64), lds, stream, g, b, active, lprof, z1, s1);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

template <int W>
int launch_ee_apply(const GridDev& g, const BatchDev& b, const int* active, const double* lprof, const double* z1,
                    const double* s1, const double* v, double* out, double* outT, hipStream_t stream) {
  constexpr int WP = (W + 1) & ~1;
  const size_t lds = (size_t)(3 * g.n + 2 * W * WP + 2 * LBUF) * sizeof(double) + (size_t)(g.n + 1) * sizeof(int) +
                     (size_t)g.n * sizeof(short);
  if (lds > 64 * 1024 || g.max_row > LBUF) return BZ_EINVAL;
  hipLaunchKernelGGL((k_ee_precond_apply<W>), dim3(b.B), dim3(64), lds, stream, g, b, active, lprof, z1, s1, v, out,
                     outT);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

template <int W, int R>
int launch_newton(const GridDev& g, const BatchDev& b, hipStream_t stream, const EEDev* ee = nullptr) {
  const size_t lds = newton_lds_bytes<W>(g.n, g.npr);
  if (lds > 64 * 1024) return BZ_EINVAL;
  const long work = (long)b.B * g.P;
  const int blocks = (int)std::min<long>((work + 255) / 256, 8192);
  hipLaunchKernelGGL(k_prepare0d, dim3(blocks), dim3(256), 0, stream, g, b);
  BZ_LAUNCH_CHECK();
  if (ee) {
    hipLaunchKernelGGL(k_ee_profile_add, dim3(blocks), dim3(256), 0, stream, g, b, *ee);
    BZ_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_newton0d<W, R>), dim3(b.B), dim3(64), lds, stream, g, b);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

// One backward-Euler step of every instance (bte_0d3v_batched.solve "transient" and
// newton_solver_batched, bte_0d3v_batched.py:59-126, 1068-1139):
//   G(x) = (x - u) - c P (M^-1 K x - mu(x) x),  P = I - uhat uhat^T,  c = dt tau,
// with x - u kept orthogonal to uhat.  The reference's Newton iteration freezes the
// Jacobian I - c Q^T (M^-1 K - mu(u) I) Q at the start of the step and takes full steps.
// In the original basis that step is
//   (K - mu' M) z_G = M G,  (K - mu' M) z_u = M uhat,  mu' = mu(u) + 1/c,
//   x <- x + (z_G - gamma z_u) / c,  gamma = uhat.z_G / uhat.z_u,
// so every iteration factorises the same matrix K - mu' M with the two right-hand
// sides, and the iterates are the reference's.  The iteration count and the stopping
// test follow the reference loop (the convergence test of an iterate is made before
// its update, and the update is still applied), per instance.
template <int W, int R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W >= 17 ? 1 : 2))) void k_transient0d(GridDev g, BatchDev b,
                                                                                           double cdt,
                                                                                           const double* ustart,
                                                                                           double* avg) {
  extern __shared__ double lds[];
  const int n = g.n, lane = threadIdx.x, inst = blockIdx.x;
  constexpr int WP = (W + 1) & ~1;
  constexpr int PIVN = (W + 3) & ~1;
  const int nchunk = (n + W - 1) / W;
  double* X = lds;
  double* Z1 = X + n;
  double* Z2 = Z1 + n;
  double* AV = Z2 + n + zpad(n, g.npr);
  double* STG = AV + n;
  double* PIV = STG + 2 * W * WP;
  double* U0 = PIV + PIVN;
  int2* SC = reinterpret_cast<int2*>(U0 + n);
  int* PO = reinterpret_cast<int*>(SC + n);
  short* LEV = reinterpret_cast<short*>(PO + (n + 1));
  double* LCH = Z1;

  double* kprof = b.kprof + (size_t)inst * g.P;
  const double* __restrict__ mprof = g.opprof + (size_t)(g.nop - 1) * g.P;
  const double* __restrict__ UH = g.uhat;
  const double* __restrict__ WV = b.vec + (size_t)inst * 4 * n;
  double* urows = b.urows + (size_t)inst * n * WP;

  fill_rowtab(g, SC, lane);
  for (int i = lane; i < n; i += 64) {
    const size_t o = (size_t)orig_index(i, g.ns, g.npr) * b.B + inst;
    X[i] = b.x[o];                              // initial iterate
    U0[i] = ustart ? ustart[o] : b.x[o];        // state at the start of the step
  }
  for (int i = lane; i <= n; i += 64) PO[i] = g.prof_off[i];
  for (int i = lane; i < nchunk; i += 64) LEV[i] = (short)g.chunk_levels[i];
  __syncthreads();

  // frozen shift: the profile holds K; make it K - mu' M, mu' from the start state
  const double mu0 = dot(WV, U0, n, lane);
  const double mup = mu0 + 1.0 / cdt;
  for (int k = lane; k < 7 * n; k += 64) {
    const int i = k / 7, dd = k - 7 * (k / 7) - 3;
    const int c = i + dd * g.ns;
    if (c >= 0 && c < n) {
      const int e = PO[i] + c - g.first_col[i];
      kprof[e] = fma(-mup, mprof[e], kprof[e]);
    }
  }
  __syncthreads();

  // G(X) into AV; returns |G|
  auto residual = [&]() -> double {
    for (int e = lane; e < 4 * g.npr; e += 64) {
      const int q = e >> 2, dd = e & 3;
      const double v = g.mchol[(size_t)(q * g.ns) * 4 + dd];
      LCH[e] = (dd == 0) ? 1.0 / v : ((q - dd >= 0) ? v : 0.0);
    }
    for (int i = lane; i < n; i += 64) AV[i] = 0.0;
    __syncthreads();
    flat_matvec(g.P, kprof, g.ent_rc, X, AV, lane);
    __syncthreads();
    mass_solve(g.ns, g.npr, LCH, AV, lane);
    __syncthreads();
    const double mu = dot(WV, X, n, lane);
    double su = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double a = fma(mup, X[i], AV[i]) - mu * X[i];      // M^-1 K x - mu x
      AV[i] = a;
      su = fma(UH[i], a, su);
    }
    su = wave_sum(su);
    double s2 = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double gi = (X[i] - U0[i]) - cdt * (AV[i] - su * UH[i]);
      AV[i] = gi;
      s2 = fma(gi, gi, s2);
    }
    __syncthreads();
    return sqrt(wave_sum(s2));
  };

  const double r0 = residual();
  double rr = r0;
  bool conv = (r0 < b.atol);
  int it = 0;
  while (!conv && it < b.max_iter) {
    if (it > 0) rr = residual();
    conv = (rr / r0 < b.rtol) || (rr < b.atol);
    for (int i = lane; i < n; i += 64) Z1[i] = g.muhat[i];
    mass_matvec(n, g.ns, mprof, g.first_col, g.prof_off, AV, Z2, lane);
    __syncthreads();
    factor_forward<W, R>(g, kprof, SC, LEV, Z1, Z2, PIV, STG, urows, WP, lane);
    __syncthreads();
    back_substitute<W>(n, urows, WP, Z1, Z2, STG, lane);
    __syncthreads();
    const double gam = dot(UH, Z2, n, lane) / dot(UH, Z1, n, lane);
    const double ic = 1.0 / cdt;
    for (int i = lane; i < n; i += 64) X[i] = fma(ic, Z2[i] - gam * Z1[i], X[i]);
    __syncthreads();
    ++it;
  }

  for (int i = lane; i < n; i += 64) {
    const size_t o = (size_t)orig_index(i, g.ns, g.npr) * b.B + inst;
    b.x[o] = X[i];
    if (avg) avg[o] += U0[i] + X[i];
  }
  if (lane == 0) {
    b.iters[inst] = it;
    b.status[inst] = conv ? 0 : 1;
    b.res_abs[inst] = rr;
    b.res0[inst] = r0;
  }
}

template <int W, int R>
int launch_transient(const GridDev& g, const BatchDev& b, double cdt, const double* ustart, double* avg,
                     hipStream_t stream) {
  const size_t lds = newton_lds_bytes<W>(g.n, g.npr) + (size_t)g.n * sizeof(double);
  if (lds > 64 * 1024) return BZ_EINVAL;
  const long work = (long)b.B * g.P;
  const int blocks = (int)std::min<long>((work + 255) / 256, 8192);
  hipLaunchKernelGGL(k_prepare0d, dim3(blocks), dim3(256), 0, stream, g, b);
  BZ_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_transient0d<W, R>), dim3(b.B), dim3(64), lds, stream, g, b, cdt, ustart, avg);
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

// Copy the C-ABI structures into the kernel argument structures and check what every
// solver kernel relies on.
int to_dev(const bz_grid0d_t* gp, const bz_batch0d_t* bp, GridDev& g, BatchDev& b) {
  g.n = gp->n; g.ns = gp->ns; g.npr = gp->npr; g.bw = gp->bw; g.S = gp->S; g.nop = gp->nop; g.P = gp->P;
  g.nslots = gp->nslots;
  g.max_row = gp->max_row;
  g.first_col = gp->first_col; g.prof_off = gp->prof_off; g.opprof = gp->opprof; g.slot_first = gp->slot_first;
  g.sched_next = gp->sched_next; g.muhat = gp->muhat; g.uhat = gp->uhat; g.u = gp->u; g.wsp = gp->wsp;
  g.mchol = gp->mchol; g.chunk_levels = gp->chunk_levels; g.ent_rc = gp->ent_rc;
  if (g.n != g.ns * g.npr || g.n < 8 || g.n > 32767 || g.nop != g.S + 3 || g.nop > kMaxOps || g.S < 1)
    return BZ_EINVAL;
  if (g.nslots < 64 || g.nslots % 64 != 0) return BZ_EINVAL;
  if (g.max_row < 1 || g.max_row > g.n) return BZ_EINVAL;
  b.B = bp->B; b.n0 = bp->n0; b.Tg = bp->Tg; b.E = bp->E; b.ns_by_n0 = bp->ns_by_n0; b.x = bp->x;
  b.kprof = bp->kprof; b.urows = bp->urows; b.vec = bp->vec; b.iters = bp->iters; b.status = bp->status;
  b.res_abs = bp->res_abs; b.res0 = bp->res0; b.atol = bp->atol; b.rtol = bp->rtol; b.alpha_min = bp->alpha_min;
  b.alpha_rho = bp->alpha_rho; b.ls_c = bp->ls_c; b.max_iter = bp->max_iter; b.mode = bp->mode;
  b.tg_scale = bp->tg_scale;
  b.tol_abs = bp->tol_abs;
  if (b.B < 1 || b.max_iter < 0) return BZ_EINVAL;
  if (!(b.atol >= 0.0) || !(b.rtol >= 0.0)) return BZ_EINVAL;
  return BZ_OK;
}

// Window widths W (>= bw + 1) and register slot levels R (= nslots / 64) the solver kernels
// are instantiated for; X(W, R) is expanded for each pair of the matching window.
#define BZ_FOR_WINDOW(Wneed, X)                                      \
  if ((Wneed) <= 8) {                                                \
    X(8, 1) X(8, 2) X(8, 3)                                          \
  } else if ((Wneed) <= 11) {                                        \
    X(11, 1) X(11, 2) X(11, 3)                                       \
  } else if ((Wneed) <= 12) {                                        \
    X(12, 1) X(12, 2) X(12, 3)                                       \
  } else if ((Wneed) <= 17) {                                        \
    X(17, 1) X(17, 2) X(17, 3) X(17, 4)                              \
  } else if ((Wneed) <= 20) {                                        \
    X(20, 1) X(20, 2) X(20, 3) X(20, 4)                              \
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

int bz_transient_step_0d(const bz_grid0d_t* gp, const bz_batch0d_t* bp, double dt_tau, const double* ustart,
                         double* avg, hipStream_t stream) {
  if (!gp || !bp || !(dt_tau > 0.0)) return BZ_EINVAL;
  GridDev g;
  BatchDev b;
  if (to_dev(gp, bp, g, b) != BZ_OK) return BZ_EINVAL;
  const int Wneed = g.bw + 1;
  const int R = g.nslots / 64;
#define BZ_X(W_, R_) \
  if (R == R_) return launch_transient<W_, R_>(g, b, dt_tau, ustart, avg, stream);
  BZ_FOR_WINDOW(Wneed, BZ_X)
#undef BZ_X
  return BZ_EINVAL;
}

static int steady_dispatch(const GridDev& g, const BatchDev& b, hipStream_t stream, const EEDev* ee) {
  // the backtracking loop terminates only for 0 < alpha_rho < 1 and alpha_min > 0
  if (!(b.alpha_rho > 0.0 && b.alpha_rho < 1.0) || !(b.alpha_min > 0.0)) return BZ_EINVAL;
  const int Wneed = g.bw + 1;
  const int R = g.nslots / 64;
#define BZ_X(W_, R_) \
  if (R == R_) return launch_newton<W_, R_>(g, b, stream, ee);
  BZ_FOR_WINDOW(Wneed, BZ_X)
#undef BZ_X
  return BZ_EINVAL;
}

int bz_steady_solve_0d(const bz_grid0d_t* gp, const bz_batch0d_t* bp, hipStream_t stream) {
  if (!gp || !bp) return BZ_EINVAL;
  GridDev g;
  BatchDev b;
  if (to_dev(gp, bp, g, b) != BZ_OK) return BZ_EINVAL;
  return steady_dispatch(g, b, stream, nullptr);
}

int bz_steady_solve_0d_ee(const bz_grid0d_t* gp, const bz_batch0d_t* bp, const bz_ee0d_t* ep, hipStream_t stream) {
  if (!gp || !bp || !ep) return BZ_EINVAL;
  GridDev g;
  BatchDev b;
  if (to_dev(gp, bp, g, b) != BZ_OK) return BZ_EINVAL;
  if (ep->ns != g.ns || ep->npr != g.npr || !ep->ccA || !ep->ccB || !ep->h || !ep->g || !ep->negam) return BZ_EINVAL;
  if ((ep->h_lo == nullptr) != (ep->g_lo == nullptr)) return BZ_EINVAL;
  EEDev e{ep->ccA, ep->ccB, ep->h, ep->g, ep->negam, ep->h_lo, ep->g_lo};
  return steady_dispatch(g, b, stream, &e);
}

int bz_ee_newton_residual(const bz_grid0d_t* gp, const bz_batch0d_t* bp, const bz_ee0d_t* ep, double* mu, double* F,
                          double* fnorm, hipStream_t stream) {
  if (!gp || !bp || !ep || !mu || !F || !fnorm) return BZ_EINVAL;
  GridDev g;
  BatchDev b;
  if (to_dev(gp, bp, g, b) != BZ_OK) return BZ_EINVAL;
  if (ep->ns != g.ns || ep->npr != g.npr || !ep->ccA || !ep->ccB || !ep->h || !ep->g || !ep->negam) return BZ_EINVAL;
  if ((ep->h_lo == nullptr) != (ep->g_lo == nullptr)) return BZ_EINVAL;
  EEDev e{ep->ccA, ep->ccB, ep->h, ep->g, ep->negam, ep->h_lo, ep->g_lo};
  const long work = (long)b.B * g.P;
  const int blocks = (int)std::min<long>((work + 255) / 256, 8192);
  hipLaunchKernelGGL(k_prepare0d, dim3(blocks), dim3(256), 0, stream, g, b);
  BZ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ee_profile_add, dim3(blocks), dim3(256), 0, stream, g, b, e);
  BZ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ee_mu_shift, dim3(b.B), dim3(64), 0, stream, g, b, mu);
  BZ_LAUNCH_CHECK();
  const size_t lds = (size_t)(3 * g.n + 4 * g.npr) * sizeof(double);
  if (lds > 64 * 1024) return BZ_EINVAL;
  hipLaunchKernelGGL(k_ee_apply, dim3(b.B), dim3(64), lds, stream, g, b, 0, (const int*)nullptr,
                     (const double*)nullptr, (const double*)nullptr, F, fnorm);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_ee_newton_factor(const bz_grid0d_t* gp, const bz_batch0d_t* bp, const int* active, double* lprof, double* z1,
                        double* s1, hipStream_t stream) {
  if (!gp || !bp || !lprof || !z1 || !s1) return BZ_EINVAL;
  GridDev g;
  BatchDev b;
  if (to_dev(gp, bp, g, b) != BZ_OK) return BZ_EINVAL;
  const int Wneed = g.bw + 1;
  const int R = g.nslots / 64;
#define BZ_X(W_, R_) \
  if (R == R_) return launch_ee_factor<W_, R_>(g, b, active, lprof, z1, s1, stream);
  BZ_FOR_WINDOW(Wneed, BZ_X)
#undef BZ_X
  return BZ_EINVAL;
}

int bz_ee_newton_precond_apply(const bz_grid0d_t* gp, const bz_batch0d_t* bp, const int* active, const double* lprof,
                               const double* z1, const double* s1, const double* v, double* z, double* zT,
                               hipStream_t stream) {
  if (!gp || !bp || !lprof || !z1 || !s1 || !v || !z) return BZ_EINVAL;
  GridDev g;
  BatchDev b;
  if (to_dev(gp, bp, g, b) != BZ_OK) return BZ_EINVAL;
  const int Wneed = g.bw + 1;
  if (Wneed <= 8) return launch_ee_apply<8>(g, b, active, lprof, z1, s1, v, z, zT, stream);
  if (Wneed <= 11) return launch_ee_apply<11>(g, b, active, lprof, z1, s1, v, z, zT, stream);
  if (Wneed <= 12) return launch_ee_apply<12>(g, b, active, lprof, z1, s1, v, z, zT, stream);
  if (Wneed <= 17) return launch_ee_apply<17>(g, b, active, lprof, z1, s1, v, z, zT, stream);
  if (Wneed <= 20) return launch_ee_apply<20>(g, b, active, lprof, z1, s1, v, z, zT, stream);
  return BZ_EINVAL;
}

int bz_ee_newton_jv(const bz_grid0d_t* gp, const bz_batch0d_t* bp, const int* active, const double* z,
                    const double* t2, double* out, hipStream_t stream) {
  if (!gp || !bp || !z || !t2 || !out) return BZ_EINVAL;
  GridDev g;
  BatchDev b;
  if (to_dev(gp, bp, g, b) != BZ_OK) return BZ_EINVAL;
  const size_t lds = (size_t)(3 * g.n + 4 * g.npr) * sizeof(double);
  if (lds > 64 * 1024) return BZ_EINVAL;
  hipLaunchKernelGGL(k_ee_apply, dim3(b.B), dim3(64), lds, stream, g, b, 1, active, z, t2, out, (double*)nullptr);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_solver0d_window(int bw) {
  const int Wneed = bw + 1;
  if (Wneed <= 8) return 8;
  if (Wneed <= 11) return 11;
  if (Wneed <= 12) return 12;
  if (Wneed <= 17) return 17;
  if (Wneed <= 20) return 20;
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
