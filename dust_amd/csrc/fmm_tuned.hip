// Tuned evaluation form of the FMM (placeholder until the kernels land).
#include "fmm_internal.h"

namespace dust {
int tuned_prepare(FmmHandle *h) {
  (void)h;
  set_error("tuned FMM form not built");
  return DUST_E_UNSUPPORTED;
}
int tuned_m2l_level(FmmHandle *, int) { return DUST_E_UNSUPPORTED; }
int tuned_eval(FmmHandle *) { return DUST_E_UNSUPPORTED; }
void tuned_release(FmmHandle *) {}
}  // namespace dust
