/* C ABI of libbz.so: MI355X (gfx950) kernels for the velocity-space hot path of the
 * B-spline x spherical-harmonic Boltzmann solver.
 *
 * Conventions
 *   - every pointer argument is a device pointer (hipMalloc / torch tensor memory)
 *     unless stated otherwise; the caller owns every buffer;
 *   - all arithmetic is FP64; matrices are row-major;
 *   - every entry point is asynchronous on `stream` and returns 0 on success,
 *     -1 (BZ_EINVAL) for invalid sizes/arguments (nothing launched),
 *     -2 (BZ_EHIP) when a HIP launch error is reported,
 *     -3 (BZ_ENOSPACE) when a request exceeds a fixed capacity of the library.
 *   - radial-major coefficient index p*ns + l (p: B-spline, l: spherical harmonic, m = 0).
 *
 * Each declaration names the reference interface it replaces
 * (ut-padas/boltzmann, paths relative to BESolver/python/).
 */
#ifndef BZ_H
#define BZ_H
#include <hip/hip_runtime.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ assembly */

/* out[p*npr+k] = sum_x w x^pw B_p^(da)(x) B_k^(db)(x) over the cubic B-spline knot
 * vector t (npr+4 knots), q Gauss-Legendre nodes (glx, glw on [-1,1]) per interval.
 * Replaces the spline branches of SpectralExpansionSpherical.compute_mass_matrix
 * (spec_spherical.py:226-277) and compute_advection_matix (spec_spherical.py:442-498). */
int bz_radial_bilinear(const double* t, int npr, const double* glx, const double* glw, int q, int pw, int da,
                       int db, double* out, hipStream_t stream);

/* Radial blocks of the binary electron-heavy collision operator: q0 (l = 0) and q1
 * (l >= 1), ptype 0 elastic / 1 excitation / 2 ionization / 3 attachment, cross
 * section given as a piecewise-linear table (tab_e [eV], tab_s [m^2]).
 * Replaces CollisionOpSP._LOp_eulerian_radial_only, CG branch
 * (collision_operator_spherical.py:79-225) reached through assemble_mat (:580-586). */
int bz_collision_radial(const double* t, int npr, const double* glx, const double* glw, int q, int ptype,
                        double vth, double c_gamma, double mByM, double threshold, double att_energy, double kappa,
                        const double* tab_e, const double* tab_s, int tab_len, double* q0, double* q1,
                        hipStream_t stream);

/* Gas-temperature diffusion radial block (l = 0).
 * Replaces CollisionOpSP.electron_gas_temperature (collision_operator_spherical.py:443-578). */
int bz_gas_temperature_radial(const double* t, int npr, const double* glx, const double* glw, int q, double vth,
                              double c_gamma, double kappa1, const double* tab_e, const double* tab_s, int tab_len,
                              double* out, hipStream_t stream);

/* out[p] = scale * sum_x w x^pw g(x) B_p(x); g = 1 (mode 0), sum of the tables (mode 1),
 * 1/sum of the tables (mode 2), exp(-x^2) (mode 3); tables concatenated with offsets
 * tab_off[ntab+1] (device).  Mode 3 is the Maxwellian projection of utils.function_to_basis
 * (utils.py:475-532).
 * Replaces utils.mass_op, temp_op, reaction_rates_op, mobility_op, diffusion_op
 * (utils.py:73-150, 321-395, 854-944). */
int bz_radial_moment(const double* t, int npr, const double* glx, const double* glw, int q, int pw, int mode,
                     double scale, double vth, double c_gamma, const double* tab_e, const double* tab_s,
                     const int* tab_off, int ntab, double* out, hipStream_t stream);

/* Dense (n x n, n = npr*ns) operator from radial blocks on the (l,l) diagonal:
 * block lmap[l] (-1 = zero) times scale.  Scatter step of the assembly routines
 * (collision_operator_spherical.py:136-144, 214-222; spec_spherical.py:270-275). */
int bz_scatter_diag_blocks(const double* blk, int npr, int ns, const int* lmap, double scale, double* out,
                           hipStream_t stream);

/* Dense E-field advection operator from mm1, mm2 and the l-coupling coefficients
 * cup[2*ns] (AM, AD of l = q+1) and cdn[2*ns] (BM, BD of l = q-1).
 * Replaces the coupling loop of compute_advection_matix (spec_spherical.py:480-498). */
int bz_advection_scatter(const double* mm1, const double* mm2, int npr, int ns, const double* cup,
                         const double* cdn, double scale, double* out, hipStream_t stream);

/* out[k*nx + i] = B_k(x_i) (dorder 0) or B_k'(x_i) (dorder 1), zero outside the support.
 * Replaces SpectralExpansionSpherical.Vq_r / Vdq_r (spec_spherical.py:338-383). */
int bz_basis_table(const double* t, int npr, const double* x, int nx, int dorder, double* out, hipStream_t stream);

/* Banded Cholesky (half bandwidth hb) of an npr x npr SPD block: lb[p*(hb+1)+d] = L[p,p-d].
 * And the dense inverse of L L^T (work: npr*npr doubles).
 * Replace SpectralExpansionSpherical.inverse_mass_mat (spec_spherical.py:299-336). */
int bz_band_cholesky(const double* A, int npr, int hb, double* lb, hipStream_t stream);
int bz_band_chol_inverse(const double* lb, int npr, int hb, double* inv, double* work, hipStream_t stream);

/* out = A^-1 rhs (rhs, out: npr x ncol row-major) for the banded SPD A (dense npr x npr,
 * half bandwidth hb) with its factor lb from bz_band_cholesky: a Cholesky solve per column
 * and `refine` steps of iterative refinement on the residual summed in compensated
 * arithmetic (work: 2*npr*ncol doubles).  Replaces the products hl_op = M^-1 (...),
 * gl_op = M^-1 (...) of CollisionOpSP.compute_rosenbluth_potentials_op
 * (collision_operator_spherical.py:966-967), which the reference forms with the explicit
 * inverse of spec_spherical.inverse_mass_mat. */
int bz_band_chol_solve_refined(const double* A, const double* lb, int npr, int hb, const double* rhs, int ncol,
                               double* out, double* work, int refine, hipStream_t stream);

/* ------------------------------------------------------------------ 0D batched solver */

/* Per-grid solver tables (device pointers), vectors in the permuted order
 * i' = (npr-1-p)*ns + l.  Built by the Python host (boltzmann_amd/plan0d.py). */
typedef struct {
  int n, ns, npr, bw, S, nop, P, nslots;
  const int* first_col;   /* [n]   first column of each row profile            */
  const int* prof_off;    /* [n+1] offsets of the row profiles                  */
  const double* opprof;   /* [nop][P]  C_0..C_{S-1}, C_Tg, A_E, M in profiles   */
  const int* slot_first;  /* [nslots] first row of every lane slot (-1: none)  */
  const int* sched_next;  /* [n]   next row of the same slot (-1: none)         */
  const double* muhat;    /* [n]   M uhat                                       */
  const double* uhat;     /* [n]   u / |u|                                      */
  const double* u;        /* [n]   mass vector u                                */
  const double* wsp;      /* [nop-1][n] W_m = Op_m^T M^-1 u (zero row: operator not in the growth term) */
  const double* mchol;    /* [n][4] banded Cholesky rows of M                   */
  const int* chunk_levels;    /* [ceil(n/W)] register levels alive per W-step chunk */
  const unsigned* ent_rc;     /* [P] (row << 16) | column of every profile entry    */
  int max_row;                /* longest row profile (entries)                      */
} bz_grid0d_t;

/* Per-batch parameters, state and scratch (device pointers). */
typedef struct {
  int B;
  const double* n0;       /* [B]                                               */
  const double* Tg;       /* [B]                                               */
  const double* E;        /* [B]  field [V/m]                                  */
  const double* ns_by_n0; /* [S][B]                                            */
  double* x;              /* (n, B): initial guess in, solution out (unnormalised) */
  double* kprof;          /* [B][P] scratch                                    */
  double* urows;          /* [B][n][W rounded up to even] scratch              */
  double* vec;            /* [B][12][n] scratch                                */
  int* iters;             /* [B]                                               */
  int* status;            /* [B] 0 converged, 1 iteration cap, 2 line search failed */
  double* res_abs;        /* [B] final |r|                                     */
  double* res0;           /* [B] initial |r|                                   */
  double atol, rtol, alpha_min, alpha_rho, ls_c;
  int max_iter;
  int mode;               /* 0: reference Jacobian, 1: exact Jacobian          */
  double tg_scale;        /* C_Tg coefficient is n0 * Tg * tg_scale            */
  const double* tol_abs;  /* [B] per-instance absolute tolerance; NULL: atol + rtol * |r0| */
  double* lprof;          /* [B][P] optional scratch of bz_transient_step_0d: the L factor of the
                             frozen matrix, so that it is factorised once per step (NULL: once
                             per Newton iteration); ignored by the steady solvers   */
} bz_batch0d_t;

/* Batched damped-Newton steady state of all B instances in one launch (one wavefront
 * per instance, every instance stops on its own criterion).
 * Replaces bte_0d3v_batched.steady_state_solve + get_rhs_and_jacobian + batched_inv
 * (bte_0d3v_batched.py:602-821, 856-979). */
int bz_steady_solve_0d(const bz_grid0d_t* grid, const bz_batch0d_t* batch, hipStream_t stream);

/* bz_steady_solve_0d with the per-instance scratch (batch->kprof, urows, vec: must be NULL)
 * taken from a static device array of the library, so that the call allocates nothing.
 * No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops that forbid hipMalloc).  Limit: 64 MiB per device,
 * i.e. B * (P + n * WP + 12 n) * 8 bytes (WP = window width rounded up to even); a larger B
 * returns -3 (BZ_ENOSPACE) with nothing launched.  Every call on a device shares that
 * device's array: calls that overlap on one device race. */
int bz_steady_solve_0d_noalloc(const bz_grid0d_t* grid, const bz_batch0d_t* batch, hipStream_t stream);

/* One backward-Euler step of size dt_tau (= dt * tau) for all B instances from the state
 * ustart (n, B), normalised u.x = 1 (NULL: batch->x), with batch->x the initial iterate on
 * entry and the new state on exit: the Newton iteration of newton_solver_batched with the
 * Jacobian frozen at the start of the step and full steps, batch->E the field at the end
 * of the step, batch->atol / rtol / max_iter its stopping rule (per instance); iters,
 * status (0 converged, 1 iteration cap), res_abs (last |G|) and res0 are written.  If avg
 * is not NULL, avg (n, B) += x_old + x_new (the reference's cycle average accumulator).
 * Replaces bte_0d3v_batched.solve(..., "transient") inner step and newton_solver_batched
 * (bte_0d3v_batched.py:59-126, 1068-1139) and bte_0d3v_batched.step (:981-1021). */
int bz_transient_step_0d(const bz_grid0d_t* grid, const bz_batch0d_t* batch, double dt_tau, const double* ustart,
                         double* avg, hipStream_t stream);

/* e-e operator data of one batch (device pointers): the banded trilinear tensors of the
 * grid and the current potentials and Coulomb prefactor of every instance. */
typedef struct {
  int ns, npr;
  const double* ccA;      /* [npr][7][7][ns][ns][ns]: cc_a[(p,q),(p+dk-3,l),(p+dr-3,s)]  */
  const double* ccB;      /* same layout for cc_b                                     */
  const double* h;        /* (B, n) first Rosenbluth potential of every instance       */
  const double* g;        /* (B, n) second potential                                   */
  const double* negam;    /* [B]   ne * gamma_a                                         */
  const double* h_lo;     /* (B, n) low parts of h (h + h_lo in twice FP64 precision), or NULL */
  const double* g_lo;     /* (B, n) low parts of g, or NULL                              */
} bz_ee0d_t;

/* bz_steady_solve_0d with the e-e operator at the given (frozen) potentials added to every
 * instance's operator: K + ne gamma_a [cc_a(., h) + cc_b(., g)] (M-space form of
 * cc = M^-1 (cc_a hl_op + cc_b gl_op)).  Outer iterations over the potentials converge to
 * the fixed point of the reference's e-e residual (bte_0d3v_batched.py:702-709, 882-979).
 * The grid tables must be planned with the e-e pattern (|k - p| <= 3 for all l, q). */
int bz_steady_solve_0d_ee(const bz_grid0d_t* grid, const bz_batch0d_t* batch, const bz_ee0d_t* ee,
                          hipStream_t stream);

/* Newton-Krylov pieces of the e-e steady state (bte_0d3v_batched.py:702-743, 882-979: the
 * residual and the Jacobian with the potential derivative cc x + cc_12 x, plus the growth
 * derivative).  With mu = w.x and K_x = K + ne gamma_a [cc_a(., Hx) + cc_b(., Gx)]:
 *   F(x) = P (M^-1 K_x x - mu x),   J v = P (M^-1 (K_x v + T2 v) - mu v - x (w.v)),
 * P = I - uhat uhat^T.  x (n, B) in batch->x (normalised, u.x = 1).
 *
 * bz_ee_newton_residual: builds the profiles K_x - mu M in batch->kprof from the potentials
 * in ee (of x), mu[b], F (B, n), fnorm[b] = |F_b|. */
int bz_ee_newton_residual(const bz_grid0d_t* grid, const bz_batch0d_t* batch, const bz_ee0d_t* ee, double* mu,
                          double* F, double* fnorm, hipStream_t stream);

/* Right preconditioner of the Newton system: z (B, n) orthogonal to uhat with
 * P M^-1 (K_x - mu M) z = v, from the profiles left by bz_ee_newton_residual.
 * bz_ee_newton_factor (once per Newton step): L of K_x - mu M into lprof ([B][P], the
 * profile positions of its entries), U into batch->urows, z1 ([B][n]) = (K_x - mu M)^-1 M uhat
 * and s1[b] = uhat.z1.  bz_ee_newton_precond_apply: z (B, n) and, optional, zT (n, B) with the
 * same values, from these factors (mass product, profile forward solve, back substitution).
 * Systems with active[b] == 0 are skipped (active optional). */
int bz_ee_newton_factor(const bz_grid0d_t* grid, const bz_batch0d_t* batch, const int* active, double* lprof,
                        double* z1, double* s1, hipStream_t stream);
int bz_ee_newton_precond_apply(const bz_grid0d_t* grid, const bz_batch0d_t* batch, const int* active,
                               const double* lprof, const double* z1, const double* s1, const double* v, double* z,
                               double* zT, hipStream_t stream);

/* out (B, n) = J z with t2 (B, n) = T2 z from bz_ee_jacobian_t2. */
int bz_ee_newton_jv(const bz_grid0d_t* grid, const bz_batch0d_t* batch, const int* active, const double* z,
                    const double* t2, double* out, hipStream_t stream);

/* Window width (columns) the solver uses for an upper bandwidth bw (-1: unsupported). */
int bz_solver0d_window(int bw);

/* Plans the tables of one grid for the fused solver from its dense device operators
 * (original radial-major order, n = npr * ns): ops[0..nop-1] = C_0..C_{S-1}, C_Tg, A_E, M
 * (n x n each), growth_ops[m] (m < nop-1) the operator whose W_m = G_m^T M^-1 u enters the
 * growth rate, or NULL (growth_ops itself may be NULL: no growth term), u the mass vector
 * (n), mass_radial the radial mass block (npr x npr), ee_pattern 1 to reserve the e-e
 * pattern |k - p| <= 3 in the profiles.  Fills *grid_out with device tables owned by the
 * library (release them with bz_free_grid0d(*handle_out)).  Synchronises `stream`.
 * Replaces the operator setup of bte_0d3v_batched.assemble_operators that the solve needs
 * (bte_0d3v_batched.py:421-442: Q, R and the premultiplied operators) for a caller of the
 * C ABI without the Python layer. */
int bz_plan_grid0d(int npr, int ns, int nop, const double* const* ops, const double* const* growth_ops,
                   const double* u, const double* mass_radial, int ee_pattern, bz_grid0d_t* grid_out,
                   void** handle_out, hipStream_t stream);
int bz_free_grid0d(void* handle);

/* Normalisation and QoIs (energy, mobility, diffusion, G rates) of B solutions.
 * Replaces bte_0d3v_batched.normalized_distribution + compute_QoIs
 * (bte_0d3v_batched.py:1163-1217).  out[b*(3+G) + {0,1,2,3..}]. */
int bz_qoi_0d(int n, int ns, int B, int G, const double* h, const double* mass, const double* temp,
              const double* mob, const double* dif, const double* rates, const double* E, const double* n0,
              double mm_scale, double mw0v3, double mw0v5, double c_gamma, int effective, double* hout, double* out,
              hipStream_t stream);

/* Setup helpers (no direct reference counterpart; steps inside
 * bte_0d3v_batched.assemble_operators, bte_0d3v_batched.py:421-442). */
int bz_gather_profile(const double* dense, int n, int ns, int npr, const int* first_col, const int* prof_off,
                      double scale, double* out, hipStream_t stream);
int bz_gemv(const double* A, int m, int k, int trans, const double* x, double alpha, double* y, hipStream_t stream);
int bz_permute(const double* in, int n, int ns, int npr, int inverse, double* out, hipStream_t stream);
int bz_mchol_perm(const double* lrev, int ns, int npr, double* out, hipStream_t stream);

/* ------------------------------------------------------------------ 1D3V batched sweeps */

/* Batched bidiagonal solve, one system per row: diag (bsz, N), subdiag (bsz, N-1),
 * b, out (bsz, N); lower != 0: forward substitution with subdiag below the diagonal,
 * else backward substitution with subdiag above it.
 * Replaces rawkernel.lower_bidiagonal_batched / upper_bidiagonal_batched and
 * bidiagonal_solve_batched (rawkernel.py:6-46, 389-413). */
int bz_bidiag_solve_batched(const double* diag, const double* subdiag, const double* b, double* out, int bsz, int N,
                            int lower, hipStream_t stream);

/* x-advection sweep of the 1D3V x-step: systems s = i*Nvt + j (i: eigenmode of the radial
 * x-advection operator, j: ordinate) of length Nx with bidiagonal (I + dt d_i cos(theta_j) D)
 * stored as diag (Nr*Nvt, Nx), subdiag (Nr*Nvt, Nx-1); lower_of[j] != 0 selects left-to-right
 * upwinding (D_L, coupling to x-1), else right to left (D_R, coupling to x+1); the inflow value
 * of every system is taken as zero (the boundary condition).
 * Replaces the two masked bidiagonal_solve_batched calls and the inflow zeroing of
 * glow1d_boltzmann.step_bte_x (glowdischarge_boltzmann_1d.py:1237-1256). */
int bz_xstep_bidiag_solve(const double* diag, const double* subdiag, const int* lower_of, const double* b,
                          double* out, int Nr, int Nvt, int Nx, hipStream_t stream);

/* out[r, l, x] = sum_i P[l, i] in[r, i, x] for r < nr: P (m, k), in (nr, k, nx), out (nr, m, nx).
 * With P = po2sh (ns, Nvt) or psh2o (Nvt, ns) this is ords_to_sph / sph_to_ords
 * (glowdischarge_boltzmann_1d.py:1032-1048; operators built at :330-334). */
int bz_slab_left_mult(const double* P, int m, int k, const double* in, int nr, int nx, double* out,
                      hipStream_t stream);

/* Batched banded triangular solve: diags (bsz, ndiags, N), diagonal d of row i at
 * [d, i] (coupling row i to i-d if lower, i+d otherwise), ndiags <= 8.
 * Replaces lower/upper_banded_tri_batched and banded_tri_solve_batched
 * (rawkernel.py:48-109, 415-439). */
int bz_banded_tri_solve_batched(const double* diags, const double* b, double* out, int bsz, int ndiags, int N,
                                int lower, hipStream_t stream);

/* Batched banded triangular product out = T b with the same layout.
 * Replaces lower/upper_banded_tri_batched_mv and banded_tri_mv_batched
 * (rawkernel.py:111-166, 441-465). */
int bz_banded_tri_mv_batched(const double* diags, const double* b, double* out, int bsz, int ndiags, int N, int lower,
                             hipStream_t stream);

/* Backward-Euler upwind advection sweeps with shared operators D_LtoR, D_RtoL
 * (ndiags, N), ndiags <= 8, direction per system from the sign of the speed.
 * vr: b, out (Nx*Nvt, Nr), speed -Ex[ix] cos_vt[ivt];  vt: b, out (Nx*Nr, Nvt), speed
 * Ex[ix] / vr[ivr];  x: b, out (Nr*Nvt, Nx), speed vr[ivr] cos_vt[ivt].
 * Replace adv_vr_BE_solve / adv_vt_BE_solve / adv_x_BE_solve and their wrappers
 * (rawkernel.py:169-375, 468-503). */
int bz_adv_vr_be(const double* D_LtoR, const double* D_RtoL, const double* Ex, const double* cos_vt, double dt,
                 const double* b, double* out, int ndiags, int Nr, int Nvt, int Nx, hipStream_t stream);
int bz_adv_vt_be(const double* D_LtoR, const double* D_RtoL, const double* Ex, const double* vr, double dt,
                 const double* b, double* out, int ndiags, int Nr, int Nvt, int Nx, hipStream_t stream);
int bz_adv_x_be(const double* D_LtoR, const double* D_RtoL, const double* vr, const double* cos_vt, double dt,
                const double* b, double* out, int ndiags, int Nr, int Nvt, int Nx, hipStream_t stream);

/* x = U^-1 b for a block upper-triangular BSR matrix with unit diagonal blocks (diagonal
 * and lower blocks skipped): data (nnzb, bsz, bsz), indptr (nbr+1), indices (nnzb);
 * b, out column-major (nbr*bsz, nrhs); bsz <= 256.
 * Replaces spsm_bsr_upper_tri_solve (BESolver/cuda/spsm_bsr_upper_tri_solve.cu:2-58)
 * as launched by SpSMSolver.solve (cusparse_spsolve.py:366-372). */
int bz_bsr_upper_unit_solve(const double* data, const int* indptr, const int* indices, int nbr, int nrhs, int bsz,
                            const double* b, double* out, hipStream_t stream);

/* C = A B for row-major contiguous A (m, k), B (k, n), C (m, n) on the FP64 matrix cores.
 * Replaces the dense eigenbasis products of step_bte_x, Qinv v and Q y with v (num_p, Nvt Nx)
 * (glowdischarge_boltzmann_1d.py:1220-1287). */
int bz_dgemm_nn(const double* A, int m, int k, const double* B, int n, double* C, hipStream_t stream);

/* ------------------------------------------------------------------ GMRES building blocks */

/* out[j*ostride] (+)= sum_e V[j*ldv + e] u[e] for j < nv (or its square root), two-pass
 * deterministic reduction over nblk partial sums (partial: nv*nblk doubles); out2[j] (if
 * not NULL) receives the same sums.  Dot products and norms of utils.gmres_pcr
 * (utils.py:1190, 1216-1222, 1230, 1236) batched over the Krylov basis. */
int bz_multi_dot(const double* V, long ldv, int nv, const double* u, long n, double* partial, int nblk, double* out,
                 int ostride, int accumulate, int take_sqrt, double* out2, hipStream_t stream);

/* dst = src + sign * sum_{j<nv} coef[j*cstride] V[j*ldv + .]: the Gram-Schmidt update and
 * x_n = x + V y of utils.gmres_pcr (utils.py:1218, 1229). */
int bz_vt_axpy(const double* V, long ldv, int nv, const double* coef, int cstride, double sign, const double* src,
               double* dst, long n, hipStream_t stream);

/* y = x / s[0] (device scalar; nothing written when s[0] == 0): V[:, i+1] = u / H[i+1, i]
 * (utils.py:1222-1223). */
int bz_scale_inv(const double* x, const double* s, double* y, long n, hipStream_t stream);

/* y[0..m-1] = argmin |H y - beta e_0| for the (m+1) x m Hessenberg H (row stride ldh),
 * Givens rotations; work: (m+1)*m doubles, y: 2m+1 doubles.  Replaces the
 * xp.linalg.lstsq call of utils.gmres_pcr (utils.py:1228). */
int bz_hess_lstsq(const double* H, int ldh, int m, const double* beta, double* work, double* y, hipStream_t stream);

/* GMRES over B independent systems of size n (one per 0D instance), basis
 * V[(j*B + b)*n + e], j = 0..m (m <= 64).  Arnoldi step i: w (B, n) = A M^-1 V_i; classical
 * Gram-Schmidt twice, V_{i+1}, the Givens-rotated column into R[b][i*m + 0..i], cs, sn
 * ([B][m]), g ([B][m+1], g[b][0] = |r0| set by the caller), resid[b] = |g[b][i+1]|;
 * a system stops (active[b] = 0, kdim[b] = i + 1) at resid <= target[b], breakdown or
 * i + 1 = m.  The batched form of the Arnoldi loop of utils.gmres_pcr (utils.py:1210-1240). */
int bz_arnoldi_batched(int n, int B, int m, int i, double* V, const double* w, double* R, double* cs, double* sn,
                       double* g, const double* target, int* active, int* kdim, double* resid, hipStream_t stream);

/* out (B, n) = sum_j y_j V_j with R y = g over the first kdim[b] steps (zero if kdim[b] = 0). */
int bz_gmres_combine_batched(int n, int B, int m, const double* V, const double* R, const double* g, const int* kdim,
                             double* out, hipStream_t stream);

/* ------------------------------------------------------------------ 1D3V v-space step */

/* bin[x] = argmin_i |E[x] - Ev[i]| (first minimum): the preconditioner membership of
 * glow1d_boltzmann.vspace_pc_setup (glowdischarge_boltzmann_1d.py:1429-1474). */
int bz_nearest_bin(const double* E, int nx, const double* Ev, int nb, int* bin, hipStream_t stream);

/* y[:, x] = P[bin[x]] u[:, x] for u, y (dof, nx), P (nbins, dof, dof): the per-point block
 * preconditioner Mmat_mvec of step_bte_v (glowdischarge_boltzmann_1d.py:1526-1533). */
int bz_binned_block_apply(const double* P, int dof, const int* bin, const double* u, int nx, double* y,
                          hipStream_t stream);

/* The v-step GMRES with every Arnoldi step on the device (no host synchronisation inside a
 * restart cycle, so a cycle can be captured in a HIP graph).  Replaces the dense products and
 * the per-point preconditioner of step_bte_v (glowdischarge_boltzmann_1d.py:1476-1592) and
 * the host-side least squares and residual tests of utils.gmres_pcr (utils.py:1172-1268):
 *   bz_vstep_matvec:   Y = Z - c (K0 Z + E_x (A Z)) on the FP64 matrix cores, Z, Y (dof x nx);
 *   bz_group_bins:     the points grouped by their E bin (counting sort, stable);
 *   bz_binned_apply:   Y[:, x] = P[bin(x)] Z[:, x] for all bins in one launch (P: nbins blocks);
 *   bz_gmres_cycle_start, bz_gmres_givens, bz_gmres_back_solve: the residual estimate of a
 *   cycle by Givens rotations, a device stop flag (done, kdone) and the final triangular solve.
 * `done` may be NULL for matvec / binned_apply outside a cycle; inside one, a set flag makes
 * every step a no-op. */
int bz_vstep_matvec(const double* K0, const double* A, const double* Z, const double* E, double c, int dof, int nx,
                    double* Y, const int* done, hipStream_t stream);
int bz_group_bins(const int* bin, int nx, int nbins, int* grp_start, int* grp_count, int* order, hipStream_t stream);
int bz_binned_apply(const double* P, int dof, int nbins, const int* grp_start, const int* grp_count,
                    const int* order, const double* Z, int nx, double* Y, const int* done, hipStream_t stream);
int bz_gmres_cycle_start(const double* scal, double atol, double rtol, int m, double* g, double* tol, int* done,
                         int* kdone, hipStream_t stream);
int bz_gmres_givens(double* H, int ldh, int i, double* cs, double* sn, double* g, const double* tol, int* done,
                    int* kdone, hipStream_t stream);
int bz_gmres_back_solve(const double* H, int ldh, const double* g, const int* kdone, double* y, int m,
                        hipStream_t stream);

/* y = x - c (Kx + E[x] Ax) over (dof, nx) arrays: Lmat_mvec of step_bte_v
 * (glowdischarge_boltzmann_1d.py:1520-1524) from the two operator products, c = dt tau. */
int bz_vstep_combine(const double* x, const double* Kx, const double* Ax, const double* E, double c, int dof, int nx,
                     double* y, hipStream_t stream);

/* ------------------------------------------------------------------ e-e assembly */

/* Partial moments at the points gx: kind 0: out[i*nq+j] = int_{t_i}^{min(x_j, t_(i+4))} v^m B_i
 * (0 for x_j below the support); kind 1: int_{max(x_j, t_i)}^{t_(i+4)} v^-m B_i (0 above it);
 * qn Gauss points (qx, qw) per knot piece.  Pm / Qm of
 * CollisionOpSP.compute_rosenbluth_potentials_op (collision_operator_spherical.py:819-909). */
int bz_partial_moment(const double* t, int npr, const double* gx, int nq, int m, int kind, const double* qx,
                      const double* qw, int qn, double* out, hipStream_t stream);

/* Potential kernels m1, m2 (npr x nq) of harmonic l from P(l+2), P(l+4), Q(l-1), Q(l-3)
 * (collision_operator_spherical.py:911-920). */
int bz_potential_kernels(int l, double m_ab, const double* gx, int nq, int npr, const double* P2, const double* P4,
                         const double* Q1, const double* Q3, double* m1, double* m2, hipStream_t stream);

/* C[i*ldc + j] = alpha sum_k A[i*sai + k*sak] B[k*sbk + j*sbj] (set-up products of the
 * potential operators, collision_operator_spherical.py:940-970). */
int bz_gemm_strided(int m, int n, int k, double alpha, const double* A, long sai, long sak, const double* B, long sbk,
                    long sbj, double* C, long ldc, hipStream_t stream);

/* Banded trilinear tensors cc_a, cc_b ([npr][7][7][ns][ns][ns]) from the weak-form term tables
 * (combination c = (q*ns+l)*ns+s: monomials off[c]..off[c+1]-1 with coefficients c[] and
 * 4 ints (power of v, derivative orders of B_p, B_k, B_r)), q Gauss points per interval.
 * Replaces CollisionOpSP.coulomb_collision_op_assembly (collision_operator_spherical.py:973-1050). */
int bz_cc_banded(const double* t, int npr, int ns, const double* glx, const double* glw, int q, const int* ta_off,
                 const double* ta_c, const int* ta_i, const int* tb_off, const double* tb_c, const int* tb_i,
                 double* ccA, double* ccB, hipStream_t stream);

/* Potentials of every instance: h[b, r*ns+s] = sum_k H[s][r][k] x[(k*ns+s), b] (and g with G);
 * H, G: (ns, npr, npr) blocks of hl_op, gl_op; x (n, B); h, g (B, n).  The products
 * hl_op x, gl_op x inside cc x x (bte_0d3v_batched.py:462-468, 702-703). */
int bz_ee_potentials(const double* H, const double* G, int ns, int npr, const double* x, int B, double* h, double* g,
                     hipStream_t stream);

/* bz_ee_potentials with the sums compensated: h + h_lo, g + g_lo carry the potentials to
 * twice FP64 precision (the e-e operator entries near v_max cancel to ~1e-6 of their terms,
 * so the rounding of plain FP64 potentials moves the fixed point). */
int bz_ee_potentials_dd(const double* H, const double* G, int ns, int npr, const double* x, int B, double* h,
                        double* g, double* h_lo, double* g_lo, hipStream_t stream);

/* bz_ee_potentials with h, g in the layout of x, (n, B). */
int bz_ee_potentials_cols(const double* H, const double* G, int ns, int npr, const double* x, int B, double* h,
                          double* g, hipStream_t stream);

/* Potential-derivative term of the e-e Jacobian (the cc_12 x product of
 * bte_0d3v_batched.py:734-743) in M-space: t2[b, (p,q)] = negam[b] sum cc_a[(p,q),(k,l),(r,s)]
 * hv[(r,s), b] x[(k,l), b] + (cc_b, gv); x, hv, gv (n, B) (hv, gv the potentials of the
 * direction), t2 (B, n); instances with active[b] == 0 are skipped (active optional). */
int bz_ee_jacobian_t2(int ns, int npr, int B, const double* ccA, const double* ccB, const double* x, const double* hv,
                      const double* gv, const double* negam, const int* active, double* t2, hipStream_t stream);

/* negam[b] = ne[b] gamma_a(x_b): m0 = c_m0 mass.x, kT = |c_kt top.x / m0|,
 * Lambda = c_lam kT^1.5 / sqrt(ne), gamma_a = ln(Lambda) c_gam (bte_0d3v_batched.py:652-661). */
int bz_ee_gamma(const double* x, int n, int B, const double* mass, const double* top, const double* ne, double c_m0,
                double c_kt, double c_lam, double c_gam, double* negam, hipStream_t stream);

/* ---------------------------------------------------------------------------------------
 * Plasma coupling of the 1D3V glow-discharge step (glowdischarge_boltzmann_1d.py).
 * ------------------------------------------------------------------------------------ */

/* Phases of bz_glow_fluid_step, run in this order when set. */
#define BZ_GLOW_MOMENTS 1     /* bte_to_fluid (:1050-1076): ne, Te, ki from partial moments  */
#define BZ_GLOW_STEP 2        /* field Dp(-phi) (:2221-2222) + backward-Euler ion step       */
#define BZ_GLOW_FIELD_AFTER 4 /* fluid_to_bte (:1078-1086): bsE = (V0/L) Dp(-phi)            */

/* Ion-fluid state and operators on n <= 200 Chebyshev points.  Matrices are stored
 * transposed (MT[c n + r] = M[r][c]).  u is (n, 3) row-major: ne, ni, Te (the reference's
 * Uin with Ns = 2, NT = 1).  status[0] converged (1) or not (0) for the last step,
 * status[1] its Newton iterations, status[2] a failure count that only grows, status[3] bit 0
 * a zero-pivot latch and bit 1 set when the last step took the partially pivoted inverse (J
 * not column diagonally dominant); info[0..2] the last |r|, |r|/|r0| and damping factor. */
typedef struct {
  int n;
  const double* DpT;   /* Dp, the Chebyshev first-derivative matrix                     */
  const double* DppT;  /* Dp Dp_mod, Dp_mod = Dp with rows 0 and n-1 set to zero        */
  const double* GT;    /* -Dp LpD^-1: field of a Poisson right-hand side               */
  double mu_i, Di;     /* ion mobility and diffusivity, nondimensional, uniform in x    */
  double n0;           /* gas density / np0                                             */
  double alpha;        /* np0 L^2 qe / (eps0 V0)                                        */
  double field_scale;  /* V0 / L                                                        */
  double atol, rtol;
  int max_iter;
  double* u;
  double* ki;          /* (n) ionisation rate per RF period per np0                     */
  double* E;           /* (n) nondimensional field of the last ion step                 */
  double* bsE;         /* (n) field of the electron v-step, V/m                         */
  const double* partial; /* (nsplit, 3, n) from bz_glow_moments                        */
  int nsplit;
  double rate_scale;   /* sqrt(4 pi) np0 tau / (2 (vth / c_gamma)^3)                    */
  int* status;
  double* info;
  const double* DpTile;  /* optional (may be NULL): Dp and Dp Dp_mod in the order of the   */
  const double* DppTile; /* kernel's register tiles, from bz_glow_tile_tables              */
} BzGlowFluid;

/* Length (doubles) of each table written by bz_glow_tile_tables. */
#define BZ_GLOW_TILE_LEN (512 * 80)

/* Partial moments of the ordinate state: partial[sp][m][x] = sum over the rows of split sp
 * (rows split in nsplit equal ranges) of W[m][row] v[row][x], v (rows, nx) with rows =
 * num_p Nvt and W[m][(r, j)] = w_m[r] po2sh[0][j] for w = mass, temperature and ionisation
 * rate operators: the l = 0 projection and the dot products of bte_to_fluid
 * (glowdischarge_boltzmann_1d.py:1063-1072) in one pass over v. */
int bz_glow_moments(const double* v, int rows, int nx, const double* W, int nsplit, double* partial,
                    hipStream_t stream);

/* One launch of the ion-fluid side of the Strang step (glowdischarge_boltzmann_1d.py:
 * 2214-2241): MOMENTS finishes bte_to_fluid; STEP computes the frozen field from the Poisson
 * problem with phi(-1) = bc_before and advances ni by the backward-Euler step of length h
 * with the Newton iteration of glow1d_utils.newton_solver (glow1d_utils.py:131-207) on the
 * residual and Jacobian of rhs_fluid / rhs_fluid_jacobian (:1088-1218, weak ion flux at both
 * ends); FIELD_AFTER writes the v-step field with phi(-1) = bc_after.  Replaces the host
 * orchestration of step_fluid (:1318-1372) and its numpy/cupy inverse.  One workgroup; the
 * Jacobian inverse is Gauss-Jordan with partial pivoting in registers. */
int bz_glow_fluid_step(const BzGlowFluid* g, int flags, double h, double bc_before, double bc_after,
                       hipStream_t stream);

/* Copies of Dp and Dp Dp_mod (given transposed, as in BzGlowFluid) in the order in which
 * bz_glow_fluid_step's register tiles read them for n points (90 < n <= 200), so that the
 * Jacobian is loaded with coalesced reads; each output holds BZ_GLOW_TILE_LEN doubles.  For
 * n <= 90 the step keeps its row layout and nothing is written.  Set-up, once per n. */
int bz_glow_tile_tables(int n, const double* DpT, const double* DppT, double* DpTile, double* DppTile,
                        hipStream_t stream);

/* ------------------------------------------------------------------ fully implicit 1D3V (FVM) */

/* Operators of the fully implicit 1D3V step on the finite-volume / ordinate velocity grid
 * (bte_1d3v_solver.py:401-505, 780-862, 1333-1502).  State v (Nr*Nvt, Nx): radial point r,
 * ordinate t (theta descending, cos < 0 on the first half), point x, x fastest.  All
 * pointers are device memory; tables are built once by the host. */
typedef struct {
  int Nr, Nvt, Nx;            /* radial points, ordinates (even), x points                  */
  int nvr, nvt, nx;           /* stencil widths pvr+1, pvt+1, px+1 (each <= 8)              */
  const double* cos_vt;       /* (Nvt)                                                      */
  const double* sin_vt;       /* (Nvt)                                                      */
  const double* vr;           /* (Nr)                                                       */
  const double* wq;           /* (Nvt) 2 pi w_t: the isotropic moment s = sum_t wq_t v_t   */
  const double* loss;         /* (Nr) loss coefficient of the collision operator, 4 pi sum  */
                              /* of the processes' dcs                                      */
  const int* g_ptr;           /* (Nr+1) CSR of the gain matrix G (summed over processes):   */
  const int* g_col;           /* (C f)[r, t] = (G s)[r] - loss[r] f[r, t]                   */
  const double* g_val;
  const double* dvrL;         /* (nvr, Nr) [d, r] = Dvr_LtoR[r, r-d]  (rows, for the        */
  const double* dvrR;         /* (nvr, Nr) [d, r] = Dvr_RtoL[r, r+d]   product)             */
  const double* dvtL;         /* (nvt, Nvt) [d, t] = Dvt_LtoR[t, t-d]                       */
  const double* dvtR;         /* (nvt, Nvt) [d, t] = Dvt_RtoL[t, t+d]                       */
  const double* dxL;          /* (nx, Nx) [k, i] = DpL[i, i-k]: the x shift, theta <= pi/2  */
  const double* dxR;          /* (nx, Nx) [k, i] = DpR[i, i+k]: theta > pi/2               */
  const double* adx;          /* (Nr) vr vth tau / L                                        */
  double dt_x;                /* time step of the x shift (gmres_pc_ax dt at init)           */
  double tau_n0np0;           /* tau n0 np0                                                  */
  double tau;                 /* tau                                                         */
  double eq;                  /* (1 / vth) q / m_e: scale of the advection operator          */
  double eaE;                 /* tau q / m_e / vth: E to the sweeps' field                   */
  /* preconditioner (XVTVRC) */
  const double* vr_l2r;       /* (nvr, Nr) sweep bands, [k, j] = D[j+k, j]                  */
  const double* vr_r2l;       /* (nvr, Nr) [k, j] = D[j-k, j]                               */
  const double* vt_l2r;       /* (nvt, Nvt) diag(sin) Dvt, same layout                      */
  const double* vt_r2l;
  const double* x_l2r;        /* (nx, Nx)                                                    */
  const double* x_r2l;
  const double* axvr;         /* (Nr) vth tau / L vr: speed of the x sweep                   */
  const double* ainv;         /* (Nr) 1 / (1 + c loss), c = gmres_pc_cv dt tau n0 np0      */
  const double* H;            /* (Nr, Nr) c diag(ainv) G (diag(1 + c loss) - c W G)^-1      */
  const int* h_lo;            /* (Nr) first nonzero column of each row of H                  */
  double* work;               /* scratch, 2 Nr Nx doubles                                     */
} BzFvm1d3v;

/* y = the fully implicit step operator applied to x: the x shift (I + dt_x adx cos D) x
 * minus dt tau (n0 np0 C x + E(x) eq A_E x), A_E the upwinded radial and angular advection
 * chosen by the sign of E(x) (Ep for E >= 0, En otherwise), the v-space term zeroed at the
 * inflow points (x = -1 for theta <= pi/2, x = 1 otherwise).  E: (Nx) in V/m.  Replaces
 * rhs_mvec of bte_1d3v.step (bte_1d3v_solver.py:1902-1931). */
int bz_fvm_rhs_mvec(const BzFvm1d3v* op, const double* E, double dt, const double* x, double* y,
                    hipStream_t stream);

/* y = the XVTVRC preconditioner applied to b: the collision solve (I - c C)^-1 (structured:
 * isotropic moment, H, diagonal), then backward-Euler upwind sweeps in vr, vt and x with
 * step alpha_dt, the inflow values zeroed before the x sweep.  y may not alias b.  Replaces
 * step_op_split(..., XVTVRC) (bte_1d3v_solver.py:1745-1870), whose collision stage applies
 * the dense PmatC or SpSM. */
int bz_fvm_precond_xvtvrc(const BzFvm1d3v* op, const double* E, double alpha_dt, const double* b, double* y,
                          hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* BZ_H */
