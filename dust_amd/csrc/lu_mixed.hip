// Tuned mixed-precision solve (placeholder until the kernels land).
#include "lu_internal.h"

using namespace dust;

extern "C" {
int dust_dgesv_mixed(int, int, const double *, int, double *, int, int, double, int, int *,
                     double *) {
  set_error("mixed-precision LU not built");
  return DUST_E_UNSUPPORTED;
}
int dust_dgesv_mixed_dev(int, int, const double *, int, double *, int, int, double, int, int *,
                         double *, void *) {
  set_error("mixed-precision LU not built");
  return DUST_E_UNSUPPORTED;
}
}
