// solve_linsys on the GPU: the block static/moving elimination of
// src/linsys/mod_linsys.f90:352-483 (dlaswp 373, dtrsm 385 and 399, dgemm 416,
// dgetrf 435, dlaswp 456, pivot offset 464-465, dgetrs 473) expressed with the
// device LU building blocks.
#include <algorithm>
#include <vector>

#include "lu_internal.h"

namespace dust {

// X U = B for the rows of B (m x n), U upper non-unit (n x n): the
// dtrsm('R','U','N','N') of mod_linsys.f90:399.  One thread per row for a
// 64-column diagonal block, GEMM for the rest.
__global__ void k_trsm_runn_block(const double *__restrict__ U, long ldu, double *B, long ldb,
                                  int m, int k, int jb) {
  __shared__ double sU[64][65];
  for (int q = threadIdx.x; q < jb * jb; q += blockDim.x) {
    int r = q % jb, c = q / jb;
    sU[r][c] = U[(k + r) + ldu * (k + c)];
  }
  __syncthreads();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double x[64];
  for (int c = 0; c < jb; ++c) x[c] = B[i + ldb * (k + c)];
  for (int c = 0; c < jb; ++c) {
    double s = x[c];
    for (int t = 0; t < c; ++t) s -= x[t] * sU[t][c];
    x[c] = s / sU[c][c];
  }
  for (int c = 0; c < jb; ++c) B[i + ldb * (k + c)] = x[c];
}

__global__ void k_add_offset(int *p, int n, int off) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] += off;
}

static int solve_linsys_dev(int rank, int ns, double *dA, long lda, int *dp, double *dres,
                            hipStream_t st, int *info) {
  const int nm = rank - ns;
  *info = 0;
  if (ns > 0 && nm > 0) {
    // Pss Asd (373): interchanges of the static pivots on the upper-right block
    DUST_TRY(lu_laswp_dev(dA + lda * ns, lda, nm, 0, ns, dp, st));
    // Usd = Lss^-1 Pss Asd (385)
    DUST_TRY(lu_trsm_llnu_dev(ns, nm, dA, lda, dA + lda * ns, lda, st));
    // Lds = Ads Uss^-1 (399)
    for (int k = 0; k < ns; k += 64) {
      int jb = std::min(64, ns - k);
      if (k > 0)
        DUST_TRY(lu_gemm_sub_dev(nm, jb, k, dA + ns, lda, dA + lda * k, lda, dA + ns + lda * k, lda,
                                 st));
      k_trsm_runn_block<<<ceil_div(nm, 64), 64, 0, st>>>(dA, lda, dA + ns, lda, nm, k, jb);
    }
    // Add -= Lds Usd (416)
    DUST_TRY(lu_gemm_sub_dev(nm, nm, ns, dA + ns, lda, dA + lda * ns, lda, dA + ns + lda * ns, lda,
                             st));
  }
  if (nm > 0) {
    // factor the moving block (435)
    DUST_TRY(lu_getrf_dev(nm, nm, dA + ns + lda * ns, lda, dp + ns, st, 64, info));
    if (*info > 0) return 0;
    // interchanges of the moving pivots on the lower-left block (456)
    if (ns > 0) DUST_TRY(lu_laswp_dev(dA + ns, lda, ns, 0, nm, dp + ns, st));
    // global pivot numbering (464-465)
    k_add_offset<<<ceil_div(nm, 256), 256, 0, st>>>(dp + ns, nm, ns);
  }
  // dgetrs with one right-hand side (473)
  DUST_TRY(lu_getrs_dev(rank, 1, dA, lda, dp, dres, rank, st));
  DUST_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace dust

using namespace dust;

namespace dust {
namespace {

// Host staging around solve_linsys_dev: A(lda, rank) with its static block
// factored, ipiv, b in; A, ipiv and res out.
int solve_block_host(int rank, int nstatic, double *A, int lda, int *ipiv, const double *b,
                     double *res) {
  DUST_CHECK(rank >= 1 && nstatic >= 0 && nstatic <= rank && lda >= rank && A && ipiv && b && res,
             DUST_E_ARG, "invalid solve_linsys arguments");
  TmpDev<double> dA, dr;
  TmpDev<int> dp;
  DUST_TRY(dA.alloc((size_t)rank * rank));
  DUST_TRY(dr.alloc(rank));
  DUST_TRY(dp.alloc(rank));
  DUST_HIP(hipMemcpy2D(dA.p, sizeof(double) * rank, A, sizeof(double) * lda, sizeof(double) * rank,
                       rank, hipMemcpyHostToDevice));
  DUST_HIP(hipMemcpy(dp.p, ipiv, sizeof(int) * rank, hipMemcpyHostToDevice));
  DUST_HIP(hipMemcpy(dr.p, b, sizeof(double) * rank, hipMemcpyHostToDevice));
  int info = 0;
  DUST_TRY(solve_linsys_dev(rank, nstatic, dA.p, rank, dp.p, dr.p, nullptr, &info));
  DUST_HIP(hipMemcpy2D(A, sizeof(double) * lda, dA.p, sizeof(double) * rank, sizeof(double) * rank,
                       rank, hipMemcpyDeviceToHost));
  DUST_HIP(hipMemcpy(ipiv, dp.p, sizeof(int) * rank, hipMemcpyDeviceToHost));
  DUST_HIP(hipMemcpy(res, dr.p, sizeof(double) * rank, hipMemcpyDeviceToHost));
  return info;
}

}  // namespace
}  // namespace dust

extern "C" {

int dust_solve_linsys(int rank, int nstatic, double *A, int lda, int *ipiv, const double *b,
                      double *res) {
  return dust::solve_block_host(rank, nstatic, A, lda, ipiv, b, res);
}

// solve_pressure_sys (mod_pressure_equation.f90:251-390) is the same block
// elimination on the surface-panel system A_pres: its side-band condition
// reads the main system's nstatic / nmoving, which for the surface-panel
// counts only differs where a side band is empty and the operation a no-op.
int dust_solve_pressure_sys(int n_sp, int nstatic_sp, double *A_pres, int lda, int *P_pres,
                            const double *b_pres, double *res_pres) {
  return dust::solve_block_host(n_sp, nstatic_sp, A_pres, lda, P_pres, b_pres, res_pres);
}

}  // extern "C"
