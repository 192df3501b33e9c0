// bz_steady_solve_0d with its scratch in a static device array: the entry point for
// callers that cannot allocate device memory (code under HIP graph capture, serving loops
// that forbid hipMalloc).  It allocates nothing: the per-instance scratch of the fused
// solver (profile, U rows, work vectors) is carved out of `scratch`, one copy per device,
// whose address is looked up once per device with hipGetSymbolAddress.
#include <mutex>

#include "bz_common.h"
#include "../../include/bz.h"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per device
constexpr int kMaxDevices = 64;

// Host code never uses this name as an address: scratch_base() returns the device
// address of this device's copy.
__device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_mu;
double* g_base[kMaxDevices] = {};

int scratch_base(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return BZ_EHIP;
  if (dev < 0 || dev >= kMaxDevices) return BZ_EINVAL;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_base[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return BZ_EHIP;
    g_base[dev] = static_cast<double*>(p);
  }
  *out = g_base[dev];
  return BZ_OK;
}

}  // namespace

extern "C" int bz_steady_solve_0d_noalloc(const bz_grid0d_t* gp, const bz_batch0d_t* bp, hipStream_t stream) {
  if (!gp || !bp) return BZ_EINVAL;
  if (bp->kprof || bp->urows || bp->vec) return BZ_EINVAL;     // the scratch comes from the library
  const int W = bz_solver0d_window(gp->bw);
  if (W < 0 || bp->B < 1 || gp->n < 1 || gp->P < 1) return BZ_EINVAL;
  const size_t WP = size_t((W + 1) & ~1);
  const size_t per = size_t(gp->P) + size_t(gp->n) * WP + 12 * size_t(gp->n);  // doubles per instance
  if (size_t(bp->B) * per * sizeof(double) > kScratchBytes) return BZ_ENOSPACE;
  double* base = nullptr;
  const int rc = scratch_base(&base);
  if (rc != BZ_OK) return rc;
  bz_batch0d_t b = *bp;
  b.kprof = base;
  b.urows = base + size_t(bp->B) * size_t(gp->P);
  b.vec = b.urows + size_t(bp->B) * size_t(gp->n) * WP;
  return bz_steady_solve_0d(gp, &b, stream);
}
