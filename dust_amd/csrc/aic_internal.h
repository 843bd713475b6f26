// Helpers of aic.hip shared with the system assembly (linsys_asm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "panel.h"

namespace dust {

__global__ void k_geo(int n, const int *__restrict__ nver, const double *__restrict__ ver,
                      Panel *P);

// b(:, r) = B s(:, r) for r < nrhs (<= 16), B (ldb, n) column-major; `part`
// holds rhs_split(n) * nrhs * n doubles of partial sums, added in a fixed
// order.
int rhs_split(int n);
void rhs_gemv(int n, int nrhs, const double *B, long ldb, const double *S, double *part,
              double *b, long ldbo, hipStream_t st);

}  // namespace dust
