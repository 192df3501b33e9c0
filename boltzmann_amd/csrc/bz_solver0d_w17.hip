// Instances of the fused 0D solver kernels for the window width W = 17 (all register
// levels R of this window).  One translation unit per window keeps the builds parallel.
#include "bz_solver0d.h"

namespace bz0d {

int steady_w17(const GridDev& g, const BatchDev& b, hipStream_t s, const EEDev* ee) {
  const int R = g.nslots / 64;
  if (R == 1) return launch_newton<17, 1>(g, b, s, ee);
  if (R == 2) return launch_newton<17, 2>(g, b, s, ee);
  if (R == 3) return launch_newton<17, 3>(g, b, s, ee);
  if (R == 4) return launch_newton<17, 4>(g, b, s, ee);
  return BZ_EINVAL;
}

int transient_w17(const GridDev& g, const BatchDev& b, double cdt, const double* ustart, double* avg,
                   hipStream_t s) {
  const int R = g.nslots / 64;
  if (R == 1) return launch_transient<17, 1>(g, b, cdt, ustart, avg, s);
  if (R == 2) return launch_transient<17, 2>(g, b, cdt, ustart, avg, s);
  if (R == 3) return launch_transient<17, 3>(g, b, cdt, ustart, avg, s);
  if (R == 4) return launch_transient<17, 4>(g, b, cdt, ustart, avg, s);
  return BZ_EINVAL;
}

int ee_factor_w17(const GridDev& g, const BatchDev& b, const int* active, double* lprof, double* z1, double* s1,
                   hipStream_t s) {
  const int R = g.nslots / 64;
  if (R == 1) return launch_ee_factor<17, 1>(g, b, active, lprof, z1, s1, s);
  if (R == 2) return launch_ee_factor<17, 2>(g, b, active, lprof, z1, s1, s);
  if (R == 3) return launch_ee_factor<17, 3>(g, b, active, lprof, z1, s1, s);
  if (R == 4) return launch_ee_factor<17, 4>(g, b, active, lprof, z1, s1, s);
  return BZ_EINVAL;
}

int ee_apply_w17(const GridDev& g, const BatchDev& b, const int* active, const double* lprof, const double* z1,
                  const double* s1, const double* v, double* out, double* outT, hipStream_t s) {
  return launch_ee_apply<17>(g, b, active, lprof, z1, s1, v, out, outT, s);
}

}  // namespace bz0d
