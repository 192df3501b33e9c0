// C entry points of the fused batched 0D solver, its set-up helpers and the QoI kernel.
// The solver kernels themselves live in bz_solver0d.h and are instantiated per window
// width in bz_solver0d_w<W>.hip.
#include "bz_solver0d.h"

namespace {

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
  const double* WV = b.vec + (size_t)inst * kVecRows * n;
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
  const double* WV = b.vec + (size_t)inst * kVecRows * n;
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
  b.lprof = bp->lprof;
  if (b.B < 1 || b.max_iter < 0) return BZ_EINVAL;
  if (!(b.atol >= 0.0) || !(b.rtol >= 0.0)) return BZ_EINVAL;
  return BZ_OK;
}

// Window widths W (>= bw + 1) the solver kernels are instantiated for (one translation
// unit per window, bz_solver0d_w<W>.hip).
#define BZ_FOR_WINDOW(Wneed, X) \
  if ((Wneed) <= 8) {            \
    X(8)                          \
  } else if ((Wneed) <= 11) {    \
    X(11)                         \
  } else if ((Wneed) <= 12) {    \
    X(12)                         \
  } else if ((Wneed) <= 17) {    \
    X(17)                         \
  } else if ((Wneed) <= 20) {    \
    X(20)                         \
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
#define BZ_X(W_) return transient_w##W_(g, b, dt_tau, ustart, avg, stream);
  BZ_FOR_WINDOW(g.bw + 1, BZ_X)
#undef BZ_X
  return BZ_EINVAL;
}

static int steady_dispatch(const GridDev& g, const BatchDev& b, hipStream_t stream, const EEDev* ee) {
  // the backtracking loop terminates only for 0 < alpha_rho < 1 and alpha_min > 0
  if (!(b.alpha_rho > 0.0 && b.alpha_rho < 1.0) || !(b.alpha_min > 0.0)) return BZ_EINVAL;
#define BZ_X(W_) return steady_w##W_(g, b, stream, ee);
  BZ_FOR_WINDOW(g.bw + 1, BZ_X)
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
#define BZ_X(W_) return ee_factor_w##W_(g, b, active, lprof, z1, s1, stream);
  BZ_FOR_WINDOW(g.bw + 1, BZ_X)
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
#define BZ_X(W_) return ee_apply_w##W_(g, b, active, lprof, z1, s1, v, z, zT, stream);
  BZ_FOR_WINDOW(g.bw + 1, BZ_X)
#undef BZ_X
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
