// Device-pointer LU building blocks shared by the LU translation units.
#pragma once
#include "common.h"

namespace dust {

// xGETRF semantics on device memory; *info as LAPACK.  nb is the panel width.
int lu_getrf_dev(int m, int n, double *dA, long lda, int *dipiv, hipStream_t st, int nb,
                 int *info);
// xGETRS('N') on device memory.
int lu_getrs_dev(int n, int nrhs, const double *dA, long lda, const int *dipiv, double *dB,
                 long ldb, hipStream_t st);
// C(M,N) -= A(M,K) B(K,N), FP64.
int lu_gemm_sub_dev(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                    double *C, long ldc, hipStream_t st);
// Row interchanges ipiv[k..k+jb) (1-based, relative to A) on ncol columns.
int lu_laswp_dev(double *A, long lda, int ncol, int k, int jb, const int *dipiv, hipStream_t st);
// B(m,n) <- L^-1 B with L unit lower (m x m).
int lu_trsm_llnu_dev(int m, int n, const double *L, long ldl, double *B, long ldb, hipStream_t st);

}  // namespace dust
