// Planning of the per-grid tables of the fused 0D solver (bz_grid0d_t) from the dense
// operators of one grid, for callers without the Python layer.  The same steps as the
// Python planner: the union sparsity pattern in the permuted order i' = (npr-1-p) ns + l,
// its row profiles, the lane-slot schedule of the factorisation by interval colouring
// (rows enter their slot 2W steps before their first column), the register levels per
// W-step chunk, and the floating-point tables (operator profiles, mass Cholesky rows, u,
// uhat, M uhat, growth vectors) formed by this library's kernels.  The planner owns every
// device array it allocates; bz_free_grid0d releases them.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

#include "bz_common.h"
#include "../../include/bz.h"

namespace {

constexpr int kPrefetchWindows = 2;

struct Plan0D {
  std::vector<void*> allocs;
  bz_grid0d_t grid;
};

template <typename T>
int dev_upload(Plan0D& pl, const std::vector<T>& v, T** out) {
  void* p = nullptr;
  const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
  if (hipMalloc(&p, bytes) != hipSuccess) return BZ_EHIP;
  pl.allocs.push_back(p);
  if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return BZ_EHIP;
  *out = static_cast<T*>(p);
  return BZ_OK;
}

int dev_alloc(Plan0D& pl, size_t count, double** out) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(1, count) * sizeof(double)) != hipSuccess) return BZ_EHIP;
  pl.allocs.push_back(p);
  *out = static_cast<double*>(p);
  return BZ_OK;
}

void release(Plan0D* pl) {
  for (void* p : pl->allocs) (void)hipFree(p);
  delete pl;
}

}  // namespace

extern "C" {

int bz_free_grid0d(void* handle) {
  if (!handle) return BZ_EINVAL;
  release(static_cast<Plan0D*>(handle));
  return BZ_OK;
}

int bz_plan_grid0d(int npr, int ns, int nop, const double* const* ops, const double* const* growth_ops,
                   const double* u, const double* mass_radial, int ee_pattern, bz_grid0d_t* grid_out,
                   void** handle_out, hipStream_t stream) {
  if (npr < 4 || ns < 1 || nop < 4 || !ops || !u || !mass_radial || !grid_out || !handle_out) return BZ_EINVAL;
  const int n = npr * ns;
  if (n >= 32768) return BZ_EINVAL;                 // ent_rc packs (row << 16) | column
  for (int m = 0; m < nop; ++m)
    if (!ops[m]) return BZ_EINVAL;
  const size_t nn = size_t(n) * n;

  // union pattern of the operators (original order) and, with e-e, |k - p| <= 3
  std::vector<uint8_t> mask(nn, 0);
  {
    std::vector<double> host(nn);
    if (hipStreamSynchronize(stream) != hipSuccess) return BZ_EHIP;
    for (int m = 0; m < nop; ++m) {
      if (hipMemcpy(host.data(), ops[m], nn * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return BZ_EHIP;
      for (size_t e = 0; e < nn; ++e) mask[e] |= (host[e] != 0.0);
    }
    if (ee_pattern)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          if (std::abs(i / ns - j / ns) <= 3) mask[size_t(i) * n + j] = 1;
  }
  // permuted pattern: perm[i'] = original index
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = (npr - 1 - i / ns) * ns + i % ns;
  std::vector<int> first(n), last(n);
  int bw = 0;
  for (int i = 0; i < n; ++i) {
    int f = i;
    for (int j = 0; j < n; ++j)
      if (mask[size_t(perm[i]) * n + perm[j]]) {
        f = std::min(f, j);
        bw = std::max(bw, j - i);
      }
    first[i] = f;
  }
  const int W = bz_solver0d_window(bw);
  if (W < 0) return BZ_EINVAL;
  std::vector<int> prof_off(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    last[i] = std::min(i + bw, n - 1);
    prof_off[i + 1] = prof_off[i] + (last[i] - first[i] + 1);
  }
  const int P = prof_off[n];
  int max_row = 0;
  for (int i = 0; i < n; ++i) max_row = std::max(max_row, prof_off[i + 1] - prof_off[i]);

  // interval colouring: row i holds a slot from step first_i - 2W to its pivot step i;
  // rows in order of their start take the lowest free slot id
  std::vector<int> start(n), order(n);
  for (int i = 0; i < n; ++i) start[i] = first[i] - kPrefetchWindows * W;
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return start[a] < start[b]; });
  using PI = std::pair<int, int>;
  std::priority_queue<PI, std::vector<PI>, std::greater<PI>> alive;   // (pivot step, slot)
  std::priority_queue<int, std::vector<int>, std::greater<int>> free_ids;
  std::vector<std::vector<int>> slot_rows;
  for (int i : order) {
    while (!alive.empty() && alive.top().first <= start[i]) {
      free_ids.push(alive.top().second);
      alive.pop();
    }
    int s;
    if (!free_ids.empty()) {
      s = free_ids.top();
      free_ids.pop();
    } else {
      s = int(slot_rows.size());
      slot_rows.emplace_back();
    }
    slot_rows[s].push_back(i);
    alive.push(PI(i, s));
  }
  const int used = int(slot_rows.size());
  const int R = std::max(1, (used + 63) / 64);
  const int nslots = 64 * R;
  std::vector<int> slot_first(nslots, -1), sched_next(n, -1);
  for (int s = 0; s < used; ++s) {
    slot_first[s] = slot_rows[s][0];
    for (size_t a = 0; a + 1 < slot_rows[s].size(); ++a) sched_next[slot_rows[s][a]] = slot_rows[s][a + 1];
  }
  const int nchunk = (n + W - 1) / W;
  std::vector<int> levels(nchunk, 1);
  for (int s = 0; s < used; ++s) {
    const int lev = s / 64 + 1;
    if (lev == 1) continue;
    for (int i : slot_rows[s]) {
      const int a = std::max(0, start[i]);
      for (int c = a / W; c <= i / W; ++c) levels[c] = std::max(levels[c], lev);
    }
  }
  std::vector<uint32_t> ent_rc;
  ent_rc.reserve(P);
  for (int i = 0; i < n; ++i)
    for (int c = first[i]; c <= last[i]; ++c) ent_rc.push_back((uint32_t(i) << 16) | uint32_t(c));

  Plan0D* pl = new Plan0D();
  int rc = BZ_OK;
  int *d_first = nullptr, *d_off = nullptr, *d_sf = nullptr, *d_sn = nullptr, *d_lev = nullptr;
  uint32_t* d_rc = nullptr;
  double *opprof = nullptr, *mchol = nullptr, *du = nullptr, *duhat = nullptr, *dmuhat = nullptr, *wsp = nullptr;
  double *tmp = nullptr, *v = nullptr, *lrev = nullptr, *mrev = nullptr, *lb = nullptr, *work = nullptr;
  do {
    if ((rc = dev_upload(*pl, first, &d_first)) || (rc = dev_upload(*pl, prof_off, &d_off)) ||
        (rc = dev_upload(*pl, slot_first, &d_sf)) || (rc = dev_upload(*pl, sched_next, &d_sn)) ||
        (rc = dev_upload(*pl, levels, &d_lev)) || (rc = dev_upload(*pl, ent_rc, &d_rc)))
      break;
    if ((rc = dev_alloc(*pl, size_t(nop) * P, &opprof)) || (rc = dev_alloc(*pl, size_t(n) * 4, &mchol)) ||
        (rc = dev_alloc(*pl, n, &du)) || (rc = dev_alloc(*pl, n, &duhat)) || (rc = dev_alloc(*pl, n, &dmuhat)) ||
        (rc = dev_alloc(*pl, size_t(nop - 1) * n, &wsp)) || (rc = dev_alloc(*pl, 2 * size_t(n), &tmp)) ||
        (rc = dev_alloc(*pl, n, &v)) || (rc = dev_alloc(*pl, size_t(npr) * 4, &lrev)) ||
        (rc = dev_alloc(*pl, size_t(npr) * npr, &mrev)) || (rc = dev_alloc(*pl, size_t(npr) * 4, &lb)) ||
        (rc = dev_alloc(*pl, 2 * size_t(npr) * ns, &work)))
      break;
    for (int m = 0; m < nop && rc == BZ_OK; ++m)
      rc = bz_gather_profile(ops[m], n, ns, npr, d_first, d_off, 1.0, opprof + size_t(m) * P, stream);
    if (rc) break;
    // mass Cholesky rows in the permuted order: the reversed radial block, factorised
    {
      std::vector<double> mr(size_t(npr) * npr), rv(size_t(npr) * npr);
      if (hipStreamSynchronize(stream) != hipSuccess ||
          hipMemcpy(mr.data(), mass_radial, mr.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        rc = BZ_EHIP;
        break;
      }
      for (int a = 0; a < npr; ++a)
        for (int b = 0; b < npr; ++b) rv[size_t(a) * npr + b] = mr[size_t(npr - 1 - a) * npr + (npr - 1 - b)];
      if (hipMemcpy(mrev, rv.data(), rv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        rc = BZ_EHIP;
        break;
      }
    }
    if ((rc = bz_band_cholesky(mrev, npr, 3, lrev, stream)) || (rc = bz_mchol_perm(lrev, ns, npr, mchol, stream)))
      break;
    // u, uhat = u / |u|, M uhat (permuted)
    if ((rc = bz_permute(u, n, ns, npr, 0, du, stream))) break;
    if ((rc = bz_gemv(u, 1, n, 0, u, 1.0, tmp, stream))) break;
    double uu = 0.0;
    if (hipMemcpyAsync(&uu, tmp, sizeof(double), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      rc = BZ_EHIP;
      break;
    }
    const double one = 1.0;
    double* d_one = tmp + 1;
    if (hipMemcpy(d_one, &one, sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      rc = BZ_EHIP;
      break;
    }
    double* uhat_orig = v;                                          // reused below
    if ((rc = bz_gemv(u, n, 1, 0, d_one, 1.0 / std::sqrt(uu), uhat_orig, stream)) ||
        (rc = bz_permute(uhat_orig, n, ns, npr, 0, duhat, stream)))
      break;
    double* muh = work;                                             // n doubles of work
    if ((rc = bz_gemv(ops[nop - 1], n, n, 0, uhat_orig, 1.0, muh, stream)) ||
        (rc = bz_permute(muh, n, ns, npr, 0, dmuhat, stream)))
      break;
    // v = M^-1 u: the radial block solved for the ns columns of u viewed as (npr, ns)
    if ((rc = bz_band_cholesky(mass_radial, npr, 3, lb, stream)) ||
        (rc = bz_band_chol_solve_refined(mass_radial, lb, npr, 3, u, ns, v, work, 2, stream)))
      break;
    if (hipMemsetAsync(wsp, 0, size_t(nop - 1) * n * sizeof(double), stream) != hipSuccess) {
      rc = BZ_EHIP;
      break;
    }
    for (int m = 0; m < nop - 1 && rc == BZ_OK; ++m) {
      const double* G = growth_ops ? growth_ops[m] : nullptr;
      if (!G) continue;
      if ((rc = bz_gemv(G, n, n, 1, v, 1.0, tmp, stream))) break;   // G^T v
      rc = bz_permute(tmp, n, ns, npr, 0, wsp + size_t(m) * n, stream);
    }
    if (rc) break;
    if (hipStreamSynchronize(stream) != hipSuccess) {
      rc = BZ_EHIP;
      break;
    }
  } while (false);
  if (rc != BZ_OK) {
    release(pl);
    return rc;
  }
  bz_grid0d_t& g = pl->grid;
  std::memset(&g, 0, sizeof(g));
  g.n = n;
  g.ns = ns;
  g.npr = npr;
  g.bw = bw;
  g.S = nop - 3;
  g.nop = nop;
  g.P = P;
  g.nslots = nslots;
  g.first_col = d_first;
  g.prof_off = d_off;
  g.opprof = opprof;
  g.slot_first = d_sf;
  g.sched_next = d_sn;
  g.muhat = dmuhat;
  g.uhat = duhat;
  g.u = du;
  g.wsp = wsp;
  g.mchol = mchol;
  g.chunk_levels = d_lev;
  g.ent_rc = d_rc;
  g.max_row = max_row;
  *grid_out = g;
  *handle_out = pl;
  return BZ_OK;
}

}  // extern "C"
