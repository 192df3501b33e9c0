// Batched triangular sweeps and products of the 1D3V solver (gfx950, FP64).
//
// Every kernel here works on many independent short recurrences (one per
// system: an x-line, a radial line or an angular line of the 1D3V state).  The
// systems are stored row by row, (bsz, N) with N contiguous, which the reference
// walks with one thread per system and stride-N loads.  Here one wavefront takes
// 64 systems; the columns of each chunk of CH columns are staged through LDS with
// coalesced loads (consecutive lanes read consecutive columns of one system),
// every lane then sweeps its own system from LDS, keeping the last ND-1
// solution values in registers, and the chunk is written back coalesced.
// Systems whose sweep runs right to left stage their chunks from the right end,
// so lanes with different directions share the same chunk loop.
// Products (no recurrence) use one thread per output element.
#include <algorithm>

#include "bz_common.h"

namespace {

constexpr int TS = 64;   // systems per workgroup (one wavefront)

enum Kind { K_BIDIAG = 0, K_BANDED = 1, K_ADV_VR = 2, K_ADV_VT = 3, K_ADV_X = 4, K_XSTEP = 5 };

template <int ND>
struct ChunkCols {
  static constexpr int CH = ND <= 3 ? 16 : 8;
  static constexpr int ST = CH + 1;   // LDS row stride (doubles)
};

struct SweepArgs {
  // per-system coefficients (K_BIDIAG: diag, subdiag; K_BANDED: diags)
  const double* diag;
  const double* sdiag;
  const double* diags;
  // shared advection operators (K_ADV_*) and the per-system advection speed inputs
  const double* DL;
  const double* DR;
  const double* p0;      // vr: Ex   vt: Ex   x: vr
  const double* p1;      // vr: cos  vt: vr   x: cos
  double dt;
  const double* b;
  double* out;
  int bsz, N, lower;
  int Nr, Nvt, Nx;
  const int* lower_of;   // K_XSTEP: direction of system s is lower_of[s % Nvt]
};

// advection speed and direction of system s (rawkernel.py:183-192, 249-254, 316-322)
template <int KIND>
__device__ __forceinline__ void adv_speed(const SweepArgs& a, int s, double& av, bool& fwd) {
  if (KIND == K_ADV_VR) {
    const int ix = s / a.Nvt, ivt = s % a.Nvt;
    av = -a.p0[ix] * a.p1[ivt];
    fwd = av >= 0;
  } else if (KIND == K_ADV_VT) {
    const int ix = s / a.Nr, ivr = s % a.Nr;
    av = a.p0[ix] / a.p1[ivr];
    fwd = -av >= 0;
  } else {
    const int ivr = s / a.Nvt, ivt = s % a.Nvt;
    av = a.p0[ivr] * a.p1[ivt];
    fwd = av >= 0;
  }
}

// neighbour term as the reference groups it: vr forward av*dt*s, vt/x forward s*av*dt,
// every backward sweep s*dt*av
template <int KIND>
__device__ __forceinline__ double adv_term(double s, double av, double dt, bool fwd) {
  if (!fwd) return s * dt * av;
  if (KIND == K_ADV_VR) return av * dt * s;
  return s * av * dt;
}

template <int KIND, int ND>
__global__ __launch_bounds__(TS) void k_sweep(SweepArgs a) {
  constexpr int CH = ChunkCols<ND>::CH;
  constexpr int ST = ChunkCols<ND>::ST;
  constexpr bool BIDIAG = (KIND == K_BIDIAG || KIND == K_XSTEP);
  constexpr int NC = (KIND == K_BANDED) ? ND : (BIDIAG ? 2 : 0);   // staged coefficient tiles
  __shared__ double tb[TS * ST];
  __shared__ double tc[(NC > 0 ? NC : 1) * TS * ST];
  __shared__ int fwd_s[TS];

  const int lane = threadIdx.x;
  const int sys0 = blockIdx.x * TS;
  const int s = sys0 + lane;
  const bool active = s < a.bsz;
  const int N = a.N;

  double av = 0.0;
  bool fwd = !(BIDIAG || KIND == K_BANDED) ? true : (a.lower != 0);
  if (KIND >= K_ADV_VR && KIND <= K_ADV_X && active) adv_speed<KIND>(a, s, av, fwd);
  if (KIND == K_XSTEP && active) fwd = a.lower_of[s % a.Nvt] != 0;
  fwd_s[lane] = fwd ? 1 : 0;
  __syncthreads();

  double y[ND > 1 ? ND - 1 : 1];
#pragma unroll
  for (int d = 0; d < (ND > 1 ? ND - 1 : 1); ++d) y[d] = 0.0;

  const int nchunk = (N + CH - 1) / CH;
  int p = 0;                                  // position along this lane's sweep
  for (int k = 0; k < nchunk; ++k) {
    const int len = min(CH, N - k * CH);
    // ---- stage b (and per-system coefficients) of chunk k of every system
    for (int e = lane; e < TS * CH; e += TS) {
      const int r = e / CH, o = e % CH;
      const int sr = sys0 + r;
      if (o < len && sr < a.bsz) {
        const int c = fwd_s[r] ? (k * CH + o) : (N - k * CH - len + o);
        tb[r * ST + o] = a.b[(size_t)sr * N + c];
        if (KIND == K_BANDED) {
#pragma unroll
          for (int d = 0; d < ND; ++d) tc[(d * TS + r) * ST + o] = a.diags[((size_t)sr * ND + d) * N + c];
        } else if (BIDIAG) {
          tc[r * ST + o] = a.diag[(size_t)sr * N + c];
          // coupling of row c to its upwind neighbour: subdiag[c-1] (lower) or subdiag[c] (upper)
          const int sc = fwd_s[r] ? c - 1 : c;
          tc[(TS + r) * ST + o] = (sc >= 0 && sc < N - 1) ? a.sdiag[(size_t)sr * (N - 1) + sc] : 0.0;
        }
      }
    }
    __syncthreads();
    // ---- sweep this lane's system through the chunk
    if (active) {
      for (int q = 0; q < len; ++q, ++p) {
        const int o = fwd ? q : (len - 1 - q);                 // tile column
        const int c = fwd ? (k * CH + o) : (N - k * CH - len + o);   // matrix row
        // the x-step zeroes the inflow value of every system before its sweep
        const double bi = (KIND == K_XSTEP && p == 0) ? 0.0 : tb[lane * ST + o];
        double val;
        if (BIDIAG) {
          const double sv = (p >= 1) ? tc[(TS + lane) * ST + o] * y[0] : 0.0;
          val = (bi - sv) / tc[lane * ST + o];
        } else if (KIND == K_BANDED) {
          double sv = 0.0;
#pragma unroll
          for (int d = 1; d < ND; ++d)
            if (d <= p) sv += tc[(d * TS + lane) * ST + o] * y[d - 1];
          val = (bi - sv) / tc[lane * ST + o];
        } else {
          const double* D = fwd ? a.DL : a.DR;
          double sv = 0.0;
#pragma unroll
          for (int d = 1; d < ND; ++d)
            if (d <= p) sv += D[(size_t)d * N + (fwd ? c - d : c + d)] * y[d - 1];
          val = (bi - adv_term<KIND>(sv, av, a.dt, fwd)) / (1.0 + av * a.dt * D[c]);
        }
#pragma unroll
        for (int d = (ND > 1 ? ND - 2 : 0); d >= 1; --d) y[d] = y[d - 1];
        y[0] = val;
        tb[lane * ST + o] = val;
      }
    }
    __syncthreads();
    // ---- write the chunk back
    for (int e = lane; e < TS * CH; e += TS) {
      const int r = e / CH, o = e % CH;
      const int sr = sys0 + r;
      if (o < len && sr < a.bsz) {
        const int c = fwd_s[r] ? (k * CH + o) : (N - k * CH - len + o);
        a.out[(size_t)sr * N + c] = tb[r * ST + o];
      }
    }
    __syncthreads();
  }
}

// The x-step's bidiagonal upwind solves as a parallel scan: one wavefront per system.  Along
// the sweep the solution obeys o_p = a_p o_(p-1) + b_p (a_p = -s_p / d_p, b_p = rhs_p / d_p,
// a_0 = 0 and rhs_0 = 0: the inflow value is zeroed), a composition of affine maps.  Each
// lane composes the maps of PER consecutive positions, the 64 lane maps are combined by a
// shuffle scan (6 steps), and each lane then re-runs its positions from the value entering
// its segment.  Every map has |a_p| < 1 (upwind, diagonally dominant), so the re-association
// changes the result only at rounding level.  The sequential sweep needed N dependent steps
// per system on 64 systems per wavefront; this takes N / 64 per system.
template <int PER>
__global__ __launch_bounds__(64) void k_xstep_scan(const double* __restrict__ diag, const double* __restrict__ sdiag,
                                                   const int* __restrict__ lower_of, const double* __restrict__ rhs,
                                                   double* __restrict__ out, int bsz, int Nvt, int N) {
  const int s = blockIdx.x;
  const int lane = threadIdx.x;
  if (s >= bsz) return;
  const bool fwd = lower_of[s % Nvt] != 0;
  const size_t row = (size_t)s * N;
  const size_t srow = (size_t)s * (N - 1);
  double al[PER], be[PER];
  double A = 1.0, B = 0.0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int p = lane * PER + q;                 // position along the sweep
    double a = 1.0, b = 0.0;                      // identity beyond the end
    if (p < N) {
      const int c = fwd ? p : N - 1 - p;          // matrix row
      const double inv = 1.0 / diag[row + c];
      const int sc = fwd ? c - 1 : c;             // coupling to the upwind neighbour
      const double sv = (p >= 1) ? sdiag[srow + sc] : 0.0;
      const double r = (p >= 1) ? rhs[row + c] : 0.0;
      a = -sv * inv;
      b = r * inv;
    }
    al[q] = a;
    be[q] = b;
    B = fma(a, B, b);
    A = a * A;
  }
  // inclusive scan of the lane maps: (A, B) <- (A, B) o (A_prev, B_prev)
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double Ap = __shfl_up(A, off, 64);
    const double Bp = __shfl_up(B, off, 64);
    if (lane >= off) {
      B = fma(A, Bp, B);
      A = A * Ap;
    }
  }
  // value entering this lane's segment: the end of the previous lane's (position 0 has a = 0)
  double o = __shfl_up(B, 1, 64);
  if (lane == 0) o = 0.0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int p = lane * PER + q;
    o = fma(al[q], o, be[q]);
    if (p < N) out[row + (fwd ? p : N - 1 - p)] = o;
  }
}

template <int KIND>
int launch_sweep(const SweepArgs& a, int nd, hipStream_t st) {
  const dim3 grid((a.bsz + TS - 1) / TS), block(TS);
  switch (nd) {
    case 1: hipLaunchKernelGGL((k_sweep<KIND, 1>), grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_sweep<KIND, 2>), grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_sweep<KIND, 3>), grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_sweep<KIND, 4>), grid, block, 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_sweep<KIND, 5>), grid, block, 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_sweep<KIND, 6>), grid, block, 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_sweep<KIND, 7>), grid, block, 0, st, a); break;
    case 8: hipLaunchKernelGGL((k_sweep<KIND, 8>), grid, block, 0, st, a); break;
    default: return BZ_EINVAL;
  }
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

// out[s, i] = sum_{d=0}^{ND-1} diags[s, d, i] * b[s, i -/+ d]  (rawkernel.py:111-166)
__global__ void k_banded_mv(const double* __restrict__ diags, const double* __restrict__ b, double* __restrict__ out,
                            int bsz, int nd, int N, int lower) {
  const size_t total = (size_t)bsz * N;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(e / N), i = (int)(e % N);
    const double* dg = diags + (size_t)s * nd * N;
    const double* bs = b + (size_t)s * N;
    double acc = 0.0;
    for (int d = 0; d < nd; ++d) {
      const int j = lower ? i - d : i + d;
      if (j < 0 || j >= N) break;
      acc += dg[(size_t)d * N + i] * bs[j];
    }
    out[e] = acc;
  }
}

// x = U^-1 b, U block upper triangular with unit diagonal blocks, BSR (indptr, indices,
// data[nnzb][bsz][bsz]); b and x column-major (n, nrhs).  One workgroup takes R right-hand
// sides (bsz*R threads); thread (r, q) owns row r of every block row of rhs q.
__global__ void k_bsr_upper_unit(const double* __restrict__ data, const int* __restrict__ indptr,
                                 const int* __restrict__ indices, int nbr, int nrhs, int bsz, int R,
                                 const double* __restrict__ b, double* __restrict__ x) {
  extern __shared__ double xs[];            // [R][bsz]
  const int r = threadIdx.x % bsz, q = threadIdx.x / bsz;
  const int rhs = blockIdx.x * R + q;
  const bool act = (q < R) && (rhs < nrhs);
  const size_t n = (size_t)nbr * bsz;
  for (int i = nbr - 1; i >= 0; --i) {
    double bi = act ? b[(size_t)rhs * n + (size_t)i * bsz + r] : 0.0;
    const int a0 = indptr[i], a1 = indptr[i + 1];
    for (int idx = a0; idx < a1; ++idx) {
      const int j = indices[idx];
      if (j <= i || j >= nbr) continue;      // diagonal (unit) and lower blocks are not used
      if (act) xs[q * bsz + r] = x[(size_t)rhs * n + (size_t)j * bsz + r];
      __syncthreads();
      if (act) {
        const double* A = data + (size_t)idx * bsz * bsz + (size_t)r * bsz;
        double sacc = 0.0;
        for (int c = 0; c < bsz; ++c) sacc += A[c] * xs[q * bsz + c];
        bi -= sacc;
      }
      __syncthreads();
    }
    if (act) x[(size_t)rhs * n + (size_t)i * bsz + r] = bi;
    __syncthreads();
  }
}

// out[r, l, x] = sum_i P[l, i] in[r, i, x]: P (m, k) applied to every (k, nx) slab of
// in (nr, k, nx); consecutive threads take consecutive x.
__global__ void k_slab_left_mult(const double* __restrict__ P, int m, int k, const double* __restrict__ in, int nr,
                                 int nx, double* __restrict__ out) {
  const size_t total = (size_t)nr * m * nx;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(e % nx);
    const size_t rl = e / nx;
    const int l = (int)(rl % m), r = (int)(rl / m);
    const double* src = in + (size_t)r * k * nx + x;
    const double* pl = P + (size_t)l * k;
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += pl[i] * src[(size_t)i * nx];
    out[e] = acc;
  }
}

}  // namespace

extern "C" {

int bz_slab_left_mult(const double* P, int m, int k, const double* in, int nr, int nx, double* out,
                      hipStream_t stream) {
  if (m < 1 || k < 1 || nr < 0 || nx < 0 || !P || !in || !out) return BZ_EINVAL;
  const size_t total = (size_t)nr * m * nx;
  if (total == 0) return BZ_OK;
  const int blocks = (int)std::min<size_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(k_slab_left_mult, dim3(blocks), dim3(256), 0, stream, P, m, k, in, nr, nx, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_bidiag_solve_batched(const double* diag, const double* subdiag, const double* b, double* out, int bsz, int N,
                            int lower, hipStream_t stream) {
  if (bsz < 0 || N < 1 || !diag || !b || !out || (N > 1 && !subdiag)) return BZ_EINVAL;
  if (bsz == 0) return BZ_OK;
  SweepArgs a{};
  a.diag = diag; a.sdiag = subdiag; a.b = b; a.out = out; a.bsz = bsz; a.N = N; a.lower = lower;
  return launch_sweep<K_BIDIAG>(a, 2, stream);
}

int bz_xstep_bidiag_solve(const double* diag, const double* subdiag, const int* lower_of, const double* b,
                          double* out, int Nr, int Nvt, int Nx, hipStream_t stream) {
  if (Nr < 1 || Nvt < 1 || Nx < 2 || !diag || !subdiag || !lower_of || !b || !out) return BZ_EINVAL;
  SweepArgs a{};
  a.diag = diag; a.sdiag = subdiag; a.lower_of = lower_of; a.b = b; a.out = out;
  a.bsz = Nr * Nvt; a.N = Nx; a.Nr = Nr; a.Nvt = Nvt; a.Nx = Nx;
  const int bsz = Nr * Nvt;
  const int per = (Nx + 63) / 64;
  if (per > 16) return launch_sweep<K_XSTEP>(a, 2, stream);       // long lines: the chunked sweep
  const dim3 grid(bsz), block(64);
  if (per <= 2) hipLaunchKernelGGL((k_xstep_scan<2>), grid, block, 0, stream, diag, subdiag, lower_of, b, out, bsz, Nvt, Nx);
  else if (per <= 4) hipLaunchKernelGGL((k_xstep_scan<4>), grid, block, 0, stream, diag, subdiag, lower_of, b, out, bsz, Nvt, Nx);
  else if (per <= 8) hipLaunchKernelGGL((k_xstep_scan<8>), grid, block, 0, stream, diag, subdiag, lower_of, b, out, bsz, Nvt, Nx);
  else hipLaunchKernelGGL((k_xstep_scan<16>), grid, block, 0, stream, diag, subdiag, lower_of, b, out, bsz, Nvt, Nx);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_banded_tri_solve_batched(const double* diags, const double* b, double* out, int bsz, int ndiags, int N,
                                int lower, hipStream_t stream) {
  if (bsz < 0 || N < 1 || ndiags < 1 || ndiags > 8 || !diags || !b || !out) return BZ_EINVAL;
  if (bsz == 0) return BZ_OK;
  SweepArgs a{};
  a.diags = diags; a.b = b; a.out = out; a.bsz = bsz; a.N = N; a.lower = lower;
  return launch_sweep<K_BANDED>(a, ndiags, stream);
}

int bz_banded_tri_mv_batched(const double* diags, const double* b, double* out, int bsz, int ndiags, int N, int lower,
                             hipStream_t stream) {
  if (bsz < 0 || N < 1 || ndiags < 1 || !diags || !b || !out) return BZ_EINVAL;
  const size_t total = (size_t)bsz * N;
  if (total == 0) return BZ_OK;
  const int blocks = (int)std::min<size_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(k_banded_mv, dim3(blocks), dim3(256), 0, stream, diags, b, out, bsz, ndiags, N, lower);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

static int adv_common(int kind, const double* DL, const double* DR, const double* p0, const double* p1, double dt,
                      const double* b, double* out, int ndiags, int Nr, int Nvt, int Nx, hipStream_t stream) {
  if (Nr < 1 || Nvt < 1 || Nx < 1 || ndiags < 1 || ndiags > 8 || !DL || !DR || !p0 || !p1 || !b || !out)
    return BZ_EINVAL;
  SweepArgs a{};
  a.DL = DL; a.DR = DR; a.p0 = p0; a.p1 = p1; a.dt = dt; a.b = b; a.out = out;
  a.Nr = Nr; a.Nvt = Nvt; a.Nx = Nx;
  if (kind == K_ADV_VR) { a.bsz = Nx * Nvt; a.N = Nr; return launch_sweep<K_ADV_VR>(a, ndiags, stream); }
  if (kind == K_ADV_VT) { a.bsz = Nx * Nr; a.N = Nvt; return launch_sweep<K_ADV_VT>(a, ndiags, stream); }
  a.bsz = Nr * Nvt; a.N = Nx;
  return launch_sweep<K_ADV_X>(a, ndiags, stream);
}

int bz_adv_vr_be(const double* D_LtoR, const double* D_RtoL, const double* Ex, const double* cos_vt, double dt,
                 const double* b, double* out, int ndiags, int Nr, int Nvt, int Nx, hipStream_t stream) {
  return adv_common(K_ADV_VR, D_LtoR, D_RtoL, Ex, cos_vt, dt, b, out, ndiags, Nr, Nvt, Nx, stream);
}

int bz_adv_vt_be(const double* D_LtoR, const double* D_RtoL, const double* Ex, const double* vr, double dt,
                 const double* b, double* out, int ndiags, int Nr, int Nvt, int Nx, hipStream_t stream) {
  return adv_common(K_ADV_VT, D_LtoR, D_RtoL, Ex, vr, dt, b, out, ndiags, Nr, Nvt, Nx, stream);
}

int bz_adv_x_be(const double* D_LtoR, const double* D_RtoL, const double* vr, const double* cos_vt, double dt,
                const double* b, double* out, int ndiags, int Nr, int Nvt, int Nx, hipStream_t stream) {
  return adv_common(K_ADV_X, D_LtoR, D_RtoL, vr, cos_vt, dt, b, out, ndiags, Nr, Nvt, Nx, stream);
}

int bz_bsr_upper_unit_solve(const double* data, const int* indptr, const int* indices, int nbr, int nrhs, int bsz,
                            const double* b, double* out, hipStream_t stream) {
  if (nbr < 1 || nrhs < 0 || bsz < 1 || bsz > 256 || !data || !indptr || !indices || !b || !out) return BZ_EINVAL;
  if (nrhs == 0) return BZ_OK;
  const int R = std::max(1, 256 / bsz);
  const int grid = (nrhs + R - 1) / R;
  hipLaunchKernelGGL(k_bsr_upper_unit, dim3(grid), dim3(bsz * R), (size_t)R * bsz * sizeof(double), stream, data,
                     indptr, indices, nbr, nrhs, bsz, R, b, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

}  // extern "C"
