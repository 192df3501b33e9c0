// Fused batched 0D steady-state Newton solver, FP64, one 64-lane wavefront per instance.
//
// Reference: bte_0d3v_batched.steady_state_solve / get_rhs_and_jacobian
// (BESolver/python/bte_0d3v_batched.py:602-821, 882-979).  The reference forms the
// dense Jacobian J = Q^T L Q - mu I in the (n-1)-dim basis Q of u-perp and inverts it.
// Here the same Newton step is obtained from the bordered system in the original
// B-spline basis (L = M^-1 K):
//      (K - mu M) y = c M uhat + d M x,   uhat^T y = uhat^T x,   d = w.y - mu
// (d = 0 reproduces the reference's Jacobian, which omits the derivative of the
// growth term; d free gives the exact Jacobian).  The step is computed in residual form
// (a third right-hand side carries the compensated residual), see k_newton0d.  K - mu M is sparse: a narrow
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
//
// This header holds the device code and the launch templates; it is included by the
// translation units that instantiate them per window width (bz_solver0d_w<W>.hip) and by
// bz_solver0d.hip (C entry points, set-up and QoI kernels).  Everything but the argument
// structures and the per-window dispatch functions is TU-local (anonymous namespace).
#pragma once
#include <algorithm>
#include "bz_common.h"
#include "../../include/bz.h"

namespace bz0d {

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
  double* lprof;            // [B][P] stored L factor (transient step), or null
};

struct EEDev {
  const double* ccA;    // [npr][7][7][ns][ns][ns]
  const double* ccB;
  const double* h;      // (B, n) original order
  const double* g;
  const double* negam;  // [B]
  const double* h_lo;   // (B, n) or null
  const double* g_lo;
};

// Per-window dispatch (one translation unit per window width W, every register level R
// of that window): steady Newton (ee: NULL or the frozen e-e potentials), transient step,
// e-e factorisation and preconditioner application.  Each returns BZ_EINVAL when no
// instance of its window matches g.nslots / 64.
#define BZ0D_DECLARE_WINDOW(W)                                                                         \
  int steady_w##W(const GridDev& g, const BatchDev& b, hipStream_t s, const EEDev* ee);               \
  int transient_w##W(const GridDev& g, const BatchDev& b, double cdt, const double* ustart, double* avg, \
                     hipStream_t s);                                                                   \
  int ee_factor_w##W(const GridDev& g, const BatchDev& b, const int* active, double* lprof, double* z1, \
                     double* s1, hipStream_t s);                                                       \
  int ee_apply_w##W(const GridDev& g, const BatchDev& b, const int* active, const double* lprof,       \
                    const double* z1, const double* s1, const double* v, double* out, double* outT,    \
                    hipStream_t s);
BZ0D_DECLARE_WINDOW(8)
BZ0D_DECLARE_WINDOW(11)
BZ0D_DECLARE_WINDOW(12)
BZ0D_DECLARE_WINDOW(17)
BZ0D_DECLARE_WINDOW(20)
#undef BZ0D_DECLARE_WINDOW

}  // namespace bz0d

namespace {

using namespace bz;
using namespace bz0d;

constexpr int kMaxOps = 16;
// rows of the per-instance work array b.vec (each n doubles): 0 growth vector w, 1 x, 2 M x,
// 3 third right-hand side, 4..10 the unshifted values of K on the band of M (row i, offset d:
// row 4 + d, entry i), 11 the Newton step
constexpr int kVecRows = 12;
// exact-Newton line search: residual evaluations per step, and the step length below which an
// accepted step counts as a stall
constexpr int kMaxTrials = 8;
constexpr double kStallAlpha = 1e-2;



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

// Dot product in compensated arithmetic (per-lane TwoProduct/TwoSum accumulation, then a
// compensated butterfly across the wavefront): about as accurate as a sum in twice FP64.
__device__ double dot_dd(const double* __restrict__ a, const double* __restrict__ b, int n, int lane) {
  double s = 0.0, c = 0.0;
  for (int i = lane; i < n; i += 64) two_sum_acc(a[i], b[i], s, c);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma clang fp contract(off)
    const double s2 = __shfl_xor(s, off, 64), c2 = __shfl_xor(c, off, 64);
    const double t = s + s2, z = t - s;
    const double e = (s - (t - z)) + (s2 - z);
    s = t;
    c = c + c2 + e;
  }
  return s + c;
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
// all a slot change needs.  b1/b2/b3 hold the row's right-hand-side values, eliminated up
// to the current step (read from Y1/Y2/Y3 when the row enters; b3 only with three
// right-hand sides).
template <int W, int R>
struct FactorState {
  int row[R], fc[R], off[R], nrow[R];
  unsigned vmask[R];
  double w[R][W], nk[R][W], b1[R], b2[R], b3[R];
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
template <int W, int R, int RA, bool SL, bool TAIL, int NRHS>
__device__ __forceinline__ void factor_chunk(int j0, int n, int bw, FactorState<W, R>& s,
                                             const glb_double* __restrict__ kprof, const lds_int* __restrict__ s_succ,
                                             lds_double* __restrict__ Y1, lds_double* __restrict__ Y2,
                                             lds_double* __restrict__ Y3,
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
          if constexpr (NRHS == 3) dst[W + 2] = s.b3[r];
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
    const double pb3 = (NRHS == 3) ? (double)s_piv[W + 2] : 0.0;
    const double rinv = fast_rcp(pv[0]);
    // --- U row j (1/U_jj first) and the eliminated rhs; every lane stores entry lane % WP
    {
      const int kk = lane % WP;
      urows[(size_t)j * WP + kk] = (kk == 0) ? rinv : s_piv[kk];
      Y1[j] = pb1;
      Y2[j] = pb2;
      if constexpr (NRHS == 3) Y3[j] = pb3;
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
      if constexpr (NRHS == 3) s.b3[r] = fma(-m, pb3, s.b3[r]);
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
          s.b1[r] = (nr >= 0) ? Y1[q] : 0.0;     // entry values of the right-hand sides of the new row
          s.b2[r] = (nr >= 0) ? Y2[q] : 0.0;
          if constexpr (NRHS == 3) s.b3[r] = (nr >= 0) ? Y3[q] : 0.0;
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
// no pivoting) and forward-eliminate NRHS (2 or 3) right-hand sides whose entry values are
// read from Y1, Y2 (and Y3) when a row enters its slot.  Y*[j] are overwritten by the
// eliminated values at pivot step j, which is never before the entry of row j.  U rows
// (1/U_jj first) go to urows.  The column prefetch is branch-free so that every step issues
// a fixed number of loads and the memory-counter waits can be counted statically.
template <int W, int R, bool SL = false, int NRHS = 2>
__device__ void factor_forward(const GridDev& g, const double* __restrict__ kprof_,
                               const int2* __restrict__ succ_,
                               double* __restrict__ y1_, double* __restrict__ y2_, double* __restrict__ y3_,
                               double* __restrict__ piv_,
                               double* __restrict__ junk_, double* __restrict__ urows_, int WP, int lane,
                               double* __restrict__ lout_ = nullptr) {
  const int n = g.n, bw = g.bw;
  const glb_double* kprof = (const glb_double*)kprof_;
  glb_double* urows = (glb_double*)urows_;
  glb_double* lout = (glb_double*)lout_;
  const lds_int* s_succ = (const lds_int*)succ_;          // the row table as int pairs
  lds_double* Y1 = (lds_double*)y1_;
  lds_double* Y2 = (lds_double*)y2_;
  lds_double* Y3 = (lds_double*)y3_;
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
    s.b3[r] = 0.0;
    if constexpr (NRHS == 3) s.b3[r] = (rw >= 0) ? (double)Y3[q] : 0.0;
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
  // Y1/Y2/Y3 double as the entry values of the right-hand sides: a row entering its
  // slot during step j has index > j, and Y*[i] is only overwritten at pivot step i.
  const int nfull = n / W;
  for (int ch = 0; ch < nfull; ++ch)
    factor_chunk<W, R, R, SL, false, NRHS>(ch * W, n, bw, s, kprof, s_succ, Y1, Y2, Y3, s_piv, s_junk, urows, WP,
                                             lane, lout);
  // last partial chunk: all levels (running a level with no rows is harmless)
  if (nfull * W < n)
    factor_chunk<W, R, R, SL, true, NRHS>(nfull * W, n, bw, s, kprof, s_succ, Y1, Y2, Y3, s_piv, s_junk, urows, WP,
                                            lane, lout);
}

// Forward solve L y = b with the unit lower factor stored by factor_forward<SL = true> in the
// profile positions of its multipliers (lprof), for one right-hand side held in Y (entry
// values in, eliminated values out; in place).  The same lane slots, row table and 2W column
// prefetch as the factorisation, with the stored multipliers streamed instead of the rows of
// K - mu M and no window update: per step and level one load and one FMA.  The elimination
// order per row is the factorisation's, so the result equals the fused elimination bitwise.
template <int W, int R, bool TAIL>
__device__ __forceinline__ void forward_chunk(int j0, int n, FactorState<W, R>& s, const glb_double* __restrict__ lp,
                                              const lds_int* __restrict__ s_succ, lds_double* __restrict__ Y,
                                              lds_double* __restrict__ s_piv, lds_double* __restrict__ s_junk,
                                              int lane) {
#pragma unroll
  for (int t = 0; t < W; ++t) {
    const int j = j0 + t;
    if (TAIL && j >= n) break;
    int lpv = 0, rr = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const unsigned long long bl = __ballot(s.row[r] == j);
      if (bl) { lpv = __ffsll((long long)bl) - 1; rr = r; }
    }
    lpv = __builtin_amdgcn_readfirstlane(lpv);
    rr = __builtin_amdgcn_readfirstlane(rr);
    const int hit = (lane == lpv) ? rr : -1;
    {
      lds_double* dst = (lane == lpv) ? s_piv : s_junk + lane;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (rr == r) dst[0] = s.b1[r];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const double pb = s_piv[0];
    Y[j] = pb;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double m = (j >= s.fc[r] && s.row[r] > j) ? s.w[r][t] : 0.0;
      s.b1[r] = fma(-m, pb, s.b1[r]);
      s.w[r][t] = ((s.vmask[r] >> t) & 1u) ? s.nk[r][t] : 0.0;
    }
    if (lane == lpv) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (hit == r) {
          const int nr = s.nrow[r];
          s.row[r] = nr;
          s.vmask[r] = 0u;
          const int q = nr >= 0 ? nr : 0;
          s.b1[r] = (nr >= 0) ? Y[q] : 0.0;
          row_enter(make_int2(s_succ[2 * q], s_succ[2 * q + 1]), nr, s.fc[r], s.off[r], s.nrow[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int c = j + 2 * W;
      const bool valid = (c >= s.fc[r]) && (c < s.row[r]);
      const unsigned idx = valid ? (unsigned)(s.off[r] + c) : 0u;
      s.nk[r][t] = lp[idx];
      s.vmask[r] = (s.vmask[r] & ~(1u << t)) | ((valid ? 1u : 0u) << t);
    }
  }
}

template <int W, int R>
__device__ void forward_solve_stored(const GridDev& g, const double* __restrict__ lprof_,
                                     const int2* __restrict__ succ_, double* __restrict__ y_,
                                     double* __restrict__ piv_, double* __restrict__ junk_, int lane) {
  const int n = g.n;
  const glb_double* lp = (const glb_double*)lprof_;
  const lds_int* s_succ = (const lds_int*)succ_;
  lds_double* Y = (lds_double*)y_;
  lds_double* s_piv = (lds_double*)piv_;
  lds_double* s_junk = (lds_double*)junk_;
  FactorState<W, R> s;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int rw = g.slot_first[r * 64 + lane];
    s.row[r] = rw;
    const int q = rw >= 0 ? rw : 0;
    row_enter(make_int2(s_succ[2 * q], s_succ[2 * q + 1]), rw, s.fc[r], s.off[r], s.nrow[r]);
    s.b1[r] = (rw >= 0) ? Y[q] : 0.0;
    s.vmask[r] = 0u;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const int c0 = k, c1 = W + k;
      double v0 = 0.0, a1 = 0.0;
      if (c0 >= s.fc[r] && c0 < rw) v0 = lp[s.off[r] + c0];
      if (c1 >= s.fc[r] && c1 < rw) {
        a1 = lp[s.off[r] + c1];
        s.vmask[r] |= 1u << k;
      }
      s.w[r][k] = v0; s.nk[r][k] = a1;
    }
  }
  __builtin_amdgcn_s_waitcnt(0x0F70);
  const int nfull = n / W;
  for (int ch = 0; ch < nfull; ++ch) forward_chunk<W, R, false>(ch * W, n, s, lp, s_succ, Y, s_piv, s_junk, lane);
  if (nfull * W < n) forward_chunk<W, R, true>(nfull * W, n, s, lp, s_succ, Y, s_piv, s_junk, lane);
}

// Back substitution U z = y for the NRHS right-hand sides held in Y1/Y2(/Y3) (in place).
// U rows are staged through LDS in chunks of 2W rows: all 64 lanes load the next chunk
// (contiguous, coalesced) into registers while lanes 0..NRHS-1 (one right-hand side each)
// sweep the current chunk from LDS; the window z[] holds z_{j+1..j+W-1} at static
// positions j mod W (chunk bases are multiples of W).
template <int W, int NRHS = 2>
__device__ void back_substitute(int n, const double* __restrict__ urows, int WP, double* __restrict__ Y1,
                                double* __restrict__ Y2, double* __restrict__ Y3, double* __restrict__ stg_,
                                int lane) {
  constexpr int CH = 2 * W;
  constexpr int PER = (CH * ((W + 1) & ~1) + 63) / 64;   // doubles per lane per chunk
  const int csize = CH * WP;
  // Y1, Y2 and stg are in LDS: say so, so that the accesses are LDS instructions (counted
  // apart from the global loads of the next chunk) rather than generic ones
  typedef __attribute__((address_space(3))) double lds_double;
  lds_double* Y = (lds_double*)(lane == 0 ? Y1 : (lane == 1 || NRHS < 3) ? Y2 : Y3);
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
    if (lane < NRHS) {
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
    // compensated: on the band of M the elastic and gas-temperature terms cancel at low
    // energy, and a plain sum would round K there at the scale of the terms
    double v = 0.0, c = 0.0;
    for (int m = 0; m < nk; ++m) {
      double cm;
      if (m < S) cm = n0 * b.ns_by_n0[(size_t)m * b.B + inst];
      else if (m == S) cm = n0 * b.Tg[inst] * b.tg_scale;
      else cm = b.E[inst];
      two_sum_acc(cm, g.opprof[(size_t)m * g.P + e], v, c);
    }
    b.kprof[idx] = v + c;
  }
  const long totw = (long)b.B * g.n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < totw; idx += (long)gridDim.x * blockDim.x) {
    const int inst = (int)(idx / g.n);
    const int i = (int)(idx - (long)inst * g.n);
    const double n0 = b.n0[inst];
    double v = 0.0, c = 0.0;
    for (int m = 0; m < nk; ++m) {
      double cm;
      if (m < S) cm = n0 * b.ns_by_n0[(size_t)m * b.B + inst];
      else if (m == S) cm = n0 * b.Tg[inst] * b.tg_scale;
      else cm = b.E[inst];
      two_sum_acc(cm, g.wsp[(size_t)m * g.n + i], v, c);
    }
    v += c;
    b.vec[(size_t)inst * kVecRows * g.n + i] = v;
  }
}

// LDS: four n-vectors R0..R3 (a pad after R2 so that R1..R2 can hold the 4*npr mass
// Cholesky coefficients), the back-substitution stage STG[2W*WP], the pivot row
// PIV[PIVN], then the row table (int2 per row).
__host__ __device__ __forceinline__ int zpad(int n, int npr) { return (4 * npr > 2 * n) ? (4 * npr - 2 * n) : 0; }

template <int W>
__host__ __device__ constexpr int piv_len() { return (W + 4) & ~1; }   // U row + three right-hand sides

template <int W>
size_t newton_lds_bytes(int n, int npr) {
  constexpr int WP = (W + 1) & ~1;
  return (size_t)(4 * n + zpad(n, npr) + 2 * W * WP + piv_len<W>()) * sizeof(double) + (size_t)n * sizeof(int2);
}

// |(I - uhat uhat^T) v| for v = ca A + cx X + cp P (two passes: explicit projection, no
// |v|^2 - (u.v)^2 cancellation); P may be null (cp ignored).
__device__ double proj_norm3(const double* A, double ca, const double* X, double cx, const double* P, double cp,
                             const double* __restrict__ uh, int n, int lane) {
  double su = 0.0;
  for (int i = lane; i < n; i += 64) {
    double v = fma(ca, A[i], cx * X[i]);
    if (P) v = fma(cp, P[i], v);
    su = fma(uh[i], v, su);
  }
  su = wave_sum(su);
  double s2 = 0.0;
  for (int i = lane; i < n; i += 64) {
    double v = fma(ca, A[i], cx * X[i]);
    if (P) v = fma(cp, P[i], v);
    const double rp = v - su * uh[i];
    s2 = fma(rp, rp, s2);
  }
  return sqrt(wave_sum(s2));
}

// Fused Newton loop of one instance (one 64-lane wavefront), residual form.
//
// F(x) = M^-1 K x - mu(x) x, mu(x) = w.x, is the reference's residual before the projection
// P = I - uhat uhat^T (bte_0d3v_batched.py:779-783); with K' = K - mu M, F = M^-1 K' x.  The
// exact Newton step p with uhat.p = 0 satisfies K' p = c M uhat + d M x - K' x, d = w.p
// (the reference's Jacobian, mode 0, takes d = 0).
//  * Residual evaluation at a point x (row_pass): x is put on u.x = 1, mu = w.x (compensated
//    dot); one pass over the rows (one lane per row) sets the band of M in the profile to
//    K' = K - mu M from the unshifted values of K (saved at the first pass; the difference is
//    formed exactly and rounded once) and forms, in compensated arithmetic (TwoProduct by fma,
//    TwoSum), the right-hand side y3 = c' M uhat - K' x for the coefficient c' of the step
//    that led to x; T = M^-1 y3, and P F = -P T.  Near convergence y3 -> 0 and is computed to
//    about twice FP64 precision, so the fixed point follows the residual rather than the
//    roundoff of the factorisation.
//  * Step: K' is factorised with the three right-hand sides M uhat, M x, y3:
//    z1 = K'^-1 M uhat, z2 = K'^-1 M x, delta = K'^-1 y3 = c' z1 - x, and
//    p = delta + a z1 + d z2 with (a, d) from uhat.p = 0 and w.p = d (mode 0: d = 0),
//    c = c' + a.  Near convergence K' is nearly singular with x close to its null vector:
//    z1 ~ x / c and z2 are large.  c' tracks c, so delta stays of the size of x; the M x
//    coefficient is not carried over (d -> 0 at the solution), so no large multiple of z2
//    enters delta only to cancel against d z2.
//  * Armijo backtracking (bte_0d3v_batched.py:936-952) on the true residual |P F(x + al p)|,
//    evaluated by the same row pass; the accepted trial's evaluation is the next iteration's.
//    (With z1 and z2 nearly parallel the 2x2 system cancels: a, d are then accurate only in
//    combination, and a residual predicted from them is not; the true residual is.)
// Global scratch rows of the instance (b.vec, kVecRows x n): 0 growth vector w, 1 x (the
// base point of the line search), 2 M x, 3 y3, 4..10 unshifted K on the band of M, 11 p.
template <int W, int R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W >= 17 ? 1 : 2))) void k_newton0d(GridDev g, BatchDev b) {
  extern __shared__ double lds[];
  const int n = g.n, lane = threadIdx.x, inst = blockIdx.x;
  constexpr int WP = (W + 1) & ~1;
  constexpr int PIVN = piv_len<W>();
  double* R0 = lds;                                        // x; y3 -> delta -> p during a step
  double* R1 = R0 + n;                                     // M uhat -> z1
  double* R2 = R1 + n;                                     // M x -> z2
  double* R3 = R2 + n + zpad(n, g.npr);                    // y3 -> T = M^-1 y3
  double* STG = R3 + n;                                    // back-substitution stage
  double* PIV = STG + 2 * W * WP;                          // published pivot row
  int2* SC = reinterpret_cast<int2*>(PIV + PIVN);
  double* LCH = R1;                                        // mass Cholesky stage (R1..R2 free)

  double* kprof = b.kprof + (size_t)inst * g.P;           // K, then K - mu M on the band of M
  const double* __restrict__ mprof = g.opprof + (size_t)(g.nop - 1) * g.P;
  const double* __restrict__ UH = g.uhat;
  double* WV = b.vec + (size_t)inst * kVecRows * n;
  double* XG = WV + n;
  double* MXG = XG + n;
  double* Y3G = MXG + n;
  double* KB0 = Y3G + n;                                   // [7][n] unshifted K on the band of M
  double* PG = KB0 + 7 * n;                                // the step p
  double* urows = b.urows + (size_t)inst * n * WP;

  // ---------------- row table into LDS, initial iterate (original (n,B) layout -> permuted)
  fill_rowtab(g, SC, lane);
  for (int i = lane; i < n; i += 64) R0[i] = b.x[(size_t)orig_index(i, g.ns, g.npr) * b.B + inst];
  __syncthreads();

  // residual at the point in R0 for the coefficient cq: x normalised in place, K' in the
  // profile, y3 in R3 and Y3G, M x in MXG, T = M^-1 y3 in R3; returns |P F(x)| = |P T|
  bool first = true;
  auto evaluate = [&](double cq) -> double {
    const double ux = dot_dd(g.u, R0, n, lane);
    for (int i = lane; i < n; i += 64) R0[i] = R0[i] / ux;
    __syncthreads();
    const double mu = dot_dd(WV, R0, n, lane);
    for (int i = lane; i < n; i += 64) {
      const int2 rtb = SC[i];
      const int base = rtb.y;
      const int fc = rtb.x & 0xffff;
      const int lc = min(i + g.bw, n - 1);
      double kv[7], mv[7];
#pragma unroll
      for (int d = 0; d < 7; ++d) {
        const int c = i + (d - 3) * g.ns;
        const bool ok = (c >= 0) && (c < n);
        const int cc = ok ? c : i;                 // the diagonal: always in the row
        mv[d] = ok ? mprof[base + cc] : 0.0;
        kv[d] = first ? kprof[base + cc] : KB0[(size_t)d * n + i];
      }
      double mx = 0.0, corr = 0.0;
#pragma unroll
      for (int d = 0; d < 7; ++d) {
        const int c = i + (d - 3) * g.ns;
        if ((c >= 0) && (c < n)) {
#pragma clang fp contract(off)
          if (first) KB0[(size_t)d * n + i] = kv[d];
          mx = fma(mv[d], R0[c], mx);
          const double pm = mu * mv[d];
          const double pe = fma(mu, mv[d], -pm);
          const double hi = kv[d] - pm;                      // TwoSum(kv, -pm)
          const double z = hi - kv[d];
          const double lo = ((kv[d] - (hi - z)) + (-pm - z)) - pe;
          const double kr = hi + lo;                         // K' rounded once
          kprof[base + c] = kr;
          corr = fma(lo - (kr - hi), R0[c], corr);           // K' - kr, to the residual
        }
      }
      double s = 0.0, e = 0.0;
#pragma unroll 4
      for (int c = fc; c <= lc; ++c) two_sum_acc(-kprof[base + c], R0[c], s, e);
      e -= corr;
      two_sum_acc(cq, g.muhat[i], s, e);
      const double y3 = s + e;
      R3[i] = y3;
      Y3G[i] = y3;
      MXG[i] = mx;
    }
    first = false;
    for (int e = lane; e < 4 * g.npr; e += 64) {
      const int q = e >> 2, dd = e & 3;
      const double v = g.mchol[(size_t)(q * g.ns) * 4 + dd];
      LCH[e] = (dd == 0) ? 1.0 / v : ((q - dd >= 0) ? v : 0.0);
    }
    __syncthreads();
    mass_solve(g.ns, g.npr, LCH, R3, lane);
    __syncthreads();
    return proj_norm3(R3, 1.0, R0, 0.0, nullptr, 0.0, UH, n, lane);
  };

  double cp = 0.0;
  int stalls = 0;
  double rt = evaluate(0.0);
  const double r0 = rt;
  const double tol = b.tol_abs ? b.tol_abs[inst] : b.atol + b.rtol * r0;
  double rprev = rt;
  int it = 0, status = 1;

  for (;;) {
    // stopping test on the true residual of x.  Mode 0 follows the reference loop
    // (bte_0d3v_batched.py:926-957): its test uses |F| of the previous iterate, refreshed
    // after the step every 10th iteration.
    const double tval = (b.mode != 0 || it == 0 || ((it - 1) % 10) == 0) ? rt : rprev;
    rprev = rt;
    if (!(tval > tol)) { status = 0; break; }
    if (it >= b.max_iter) { status = 1; break; }
    // ---- right-hand sides and factorisation of K'
    for (int i = lane; i < n; i += 64) {
      XG[i] = R0[i];
      R0[i] = Y3G[i];
      R1[i] = g.muhat[i];
      R2[i] = MXG[i];
    }
    __syncthreads();
    factor_forward<W, R, false, 3>(g, kprof, SC, R1, R2, R0, PIV, STG, urows, WP, lane);
    __syncthreads();
    back_substitute<W, 3>(n, urows, WP, R1, R2, R0, STG, lane);
    __syncthreads();
    // ---- step p = delta + a z1 + b z2
    double s11 = 0.0, s12 = 0.0, s13 = 0.0, wz1 = 0.0, wz2 = 0.0, wz3 = 0.0;
    for (int i = lane; i < n; i += 64) {
      s11 = fma(UH[i], R1[i], s11);
      s12 = fma(UH[i], R2[i], s12);
      s13 = fma(UH[i], R0[i], s13);
      wz1 = fma(WV[i], R1[i], wz1);
      wz2 = fma(WV[i], R2[i], wz2);
      wz3 = fma(WV[i], R0[i], wz3);
    }
    s11 = wave_sum(s11); s12 = wave_sum(s12); s13 = wave_sum(s13);
    wz1 = wave_sum(wz1); wz2 = wave_sum(wz2); wz3 = wave_sum(wz3);
    double a, bb;
    if (b.mode == 0) {
      a = -s13 / s11;
      bb = 0.0;
    } else {
      const double s22 = wz2 - 1.0;
      const double e2 = -wz3;
      const double det = s11 * s22 - s12 * wz1;
      a = (-s13 * s22 - s12 * e2) / det;
      bb = (s11 * e2 + wz1 * s13) / det;
    }
    const double c_new = cp + a;
    for (int i = lane; i < n; i += 64) PG[i] = fma(a, R1[i], fma(bb, R2[i], R0[i]));
    __syncthreads();
    // ---- line search on the true residual at x + al p
    const double rbase = rt;
    const double ae2 = rt * rt;
    double alpha = 1.0;
    bool diverged = false;
    int trials = 0;
    for (;;) {
      for (int i = lane; i < n; i += 64) R0[i] = fma(alpha, PG[i], XG[i]);
      __syncthreads();
      const double rnew = evaluate(c_new);
      ++trials;
      if (!(0.5 * rnew * rnew > 0.5 * ae2 - b.ls_c * alpha * ae2)) {
        rt = rnew;
        break;
      }
      // exact Newton: at most kMaxTrials evaluations per step (a step that needs more does
      // not lead to a root; see the stall rule below)
      if (b.mode != 0 && trials >= kMaxTrials) { diverged = true; break; }
      if (b.mode == 0) {
        alpha *= b.alpha_rho;                                // the reference's backtracking
      } else {
        // exact Newton direction: slope of 0.5 |P F|^2 at 0 is -|P F|^2; the minimiser of the
        // quadratic through f(0), f'(0), f(alpha), kept within [alpha / 10, alpha_rho alpha]
        const double f0 = 0.5 * ae2, g0 = -ae2, f1 = 0.5 * rnew * rnew;
        const double den = 2.0 * (f1 - f0 - g0 * alpha);
        double aq = (den > 0.0) ? -g0 * alpha * alpha / den : b.alpha_rho * alpha;
        aq = fmin(fmax(aq, 0.1 * alpha), b.alpha_rho * alpha);
        alpha = (aq == aq) ? aq : b.alpha_rho * alpha;      // NaN guard
      }
      if (alpha < b.alpha_min) { diverged = true; break; }
    }
    if (diverged) {
      // back to x: its profile, y3 and M x are re-formed if the caller continues
      for (int i = lane; i < n; i += 64) R0[i] = XG[i];
      __syncthreads();
      status = 2;
      break;
    }
    cp = c_new;
    ++it;
    // exact Newton: two steps in a row accepted only at alpha < kStallAlpha, or reducing |r|
    // by less than 10 %, make a stalled iteration (measured on the 4096-instance batch: no
    // instance that converges backtracks below alpha = 0.3; the instances without a root of
    // the batched residual stall at |r| / |r0| ~ 1e-3 with such steps)
    if (b.mode != 0) {
      stalls = (alpha < kStallAlpha || rt > 0.9 * rbase) ? stalls + 1 : 0;
      if (stalls >= 2) { status = 2; break; }
    }
  }

  // ---------------- write back (original order)
  for (int i = lane; i < n; i += 64) b.x[(size_t)orig_index(i, g.ns, g.npr) * b.B + inst] = R0[i];
  if (lane == 0) {
    b.iters[inst] = it;
    b.status[inst] = status;
    b.res_abs[inst] = rt;
    b.res0[inst] = r0;
  }
}

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
  constexpr int PIVN = piv_len<W>();
  double* Z1 = lds + n;
  double* Z2 = Z1 + n;
  double* AV = Z2 + n + zpad(n, g.npr);
  double* STG = AV + n;
  double* PIV = STG + 2 * W * WP;
  int2* SC = reinterpret_cast<int2*>(PIV + PIVN);
  const double* kprof = b.kprof + (size_t)inst * g.P;
  double* urows = b.urows + (size_t)inst * n * WP;
  fill_rowtab(g, SC, lane);
  for (int i = lane; i < n; i += 64) {
    Z1[i] = g.muhat[i];
    Z2[i] = 0.0;
  }
  __syncthreads();
  factor_forward<W, R, true>(g, kprof, SC, Z1, Z2, nullptr, PIV, STG, urows, WP, lane, lprof + (size_t)inst * g.P);
  __syncthreads();
  back_substitute<W>(n, urows, WP, Z1, Z2, nullptr, STG, lane);
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
  back_substitute<W>(n, urows, WP, Y1, Y2, nullptr, STG, lane);
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
  constexpr int PIVN = piv_len<W>();
  double* X = lds;
  double* Z1 = X + n;
  double* Z2 = Z1 + n;
  double* AV = Z2 + n + zpad(n, g.npr);
  double* STG = AV + n;
  double* PIV = STG + 2 * W * WP;
  double* U0 = PIV + PIVN;
  int2* SC = reinterpret_cast<int2*>(U0 + n);
  double* LCH = Z1;

  double* kprof = b.kprof + (size_t)inst * g.P;
  const double* __restrict__ mprof = g.opprof + (size_t)(g.nop - 1) * g.P;
  const double* __restrict__ UH = g.uhat;
  const double* __restrict__ WV = b.vec + (size_t)inst * kVecRows * n;
  double* urows = b.urows + (size_t)inst * n * WP;

  fill_rowtab(g, SC, lane);
  for (int i = lane; i < n; i += 64) {
    const size_t o = (size_t)orig_index(i, g.ns, g.npr) * b.B + inst;
    X[i] = b.x[o];                              // initial iterate
    U0[i] = ustart ? ustart[o] : b.x[o];        // state at the start of the step
  }
  __syncthreads();

  // frozen shift: the profile holds K; make it K - mu' M, mu' from the start state
  const double mu0 = dot(WV, U0, n, lane);
  const double mup = mu0 + 1.0 / cdt;
  for (int k = lane; k < 7 * n; k += 64) {
    const int i = k / 7, dd = k - 7 * (k / 7) - 3;
    const int c = i + dd * g.ns;
    if (c >= 0 && c < n) {
      const int e = SC[i].y + c;
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

  // With b.lprof the frozen matrix is factorised once (its multipliers stored in the
  // instance's profile positions of lprof, z_u = (K - mu' M)^-1 M uhat kept in vec row 1);
  // later iterations forward-solve with the stored L and back-substitute one right-hand side.
  double* lp = b.lprof ? b.lprof + (size_t)inst * g.P : nullptr;
  double* ZU = b.vec + (size_t)inst * kVecRows * n + n;
  double su = 1.0;
  const double r0 = residual();
  double rr = r0;
  bool conv = (r0 < b.atol);
  int it = 0;
  while (!conv && it < b.max_iter) {
    if (it > 0) rr = residual();
    conv = (rr / r0 < b.rtol) || (rr < b.atol);
    mass_matvec(n, g.ns, mprof, g.first_col, g.prof_off, AV, Z2, lane);
    if (it == 0 || lp == nullptr) {
      for (int i = lane; i < n; i += 64) Z1[i] = g.muhat[i];
      __syncthreads();
      if (lp) factor_forward<W, R, true>(g, kprof, SC, Z1, Z2, nullptr, PIV, STG, urows, WP, lane, lp);
      else factor_forward<W, R>(g, kprof, SC, Z1, Z2, nullptr, PIV, STG, urows, WP, lane);
      __syncthreads();
      back_substitute<W>(n, urows, WP, Z1, Z2, nullptr, STG, lane);
      __syncthreads();
      su = dot(UH, Z1, n, lane);
      if (lp)
        for (int i = lane; i < n; i += 64) ZU[i] = Z1[i];
    } else {
      __syncthreads();
      forward_solve_stored<W, R>(g, lp, SC, Z2, PIV, STG, lane);
      __syncthreads();
      back_substitute<W, 1>(n, urows, WP, Z2, nullptr, nullptr, STG, lane);
      for (int i = lane; i < n; i += 64) Z1[i] = ZU[i];
      __syncthreads();
    }
    const double gam = dot(UH, Z2, n, lane) / su;
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
  const size_t lds = newton_lds_bytes<W>(g.n, g.npr) + (size_t)g.n * sizeof(double);   // + U0
  if (lds > 64 * 1024) return BZ_EINVAL;
  const long work = (long)b.B * g.P;
  const int blocks = (int)std::min<long>((work + 255) / 256, 8192);
  hipLaunchKernelGGL(k_prepare0d, dim3(blocks), dim3(256), 0, stream, g, b);
  BZ_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_transient0d<W, R>), dim3(b.B), dim3(64), lds, stream, g, b, cdt, ustart, avg);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

}  // namespace
