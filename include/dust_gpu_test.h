/* Test-only entry points of libdust_gpu.so.
 *
 * These are not part of the drop-in interface (include/dust_gpu.h) and have
 * no counterpart in the reference: they expose building blocks of the FP64
 * LU factorisation so that tests can check them one at a time.  Device
 * pointers, synchronous. */
#ifndef DUST_GPU_TEST_H
#define DUST_GPU_TEST_H

#ifdef __cplusplus
extern "C" {
#endif

/* C(M,N) -= A(M,K) B(K,N) in FP64 on the matrix cores. */
int dust_test_gemm_sub_dev(int M, int N, int K, const double *A, int lda, const double *B,
                           int ldb, double *C, int ldc);
/* B(m,n) <- L^-1 B with L unit lower (m x m). */
int dust_test_trsm_llnu_dev(int m, int n, const double *L, int ldl, double *B, int ldb);
/* Writes v into the error word of the dgetrs workspace (no-op before the
 * first dgetrs call allocates it): a stand-in for an earlier time-out. */
int dust_test_getrs_set_error_word(unsigned v);

#ifdef __cplusplus
}
#endif
#endif /* DUST_GPU_TEST_H */
