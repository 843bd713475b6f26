// Device-pointer LU building blocks shared by the LU translation units.
#pragma once
#include "common.h"

namespace dust {

// ---------------------------------------------------------------------------
// Grid barrier for cooperative launches: arrival counter + generation word.
// Spins are bounded; a timed-out barrier raises flag[2] and every block
// leaves the kernel at the next check.
// ---------------------------------------------------------------------------
// Only the arriving thread fences at device scope: the workgroup barrier has
// already made the block's stores visible in L2, and a device-scope release
// writes the whole L2 back, so one per block is enough.
__device__ inline bool grid_sync(unsigned *bar, unsigned nblocks, unsigned &gen) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    __threadfence();
    unsigned prev = atomicAdd(&bar[0], 1u);
    if (prev == nblocks - 1) {
      atomicExch(&bar[0], 0u);
      __threadfence();
      atomicAdd(&bar[1], 1u);
    } else {
      long spins = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > (1L << 26)) {
          atomicExch(&bar[2], 1u);
          ok = 0;
          break;
        }
        if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          ok = 0;
          break;
        }
      }
    }
    __threadfence();
    gen += 1;
  }
  __syncthreads();
  return ok != 0;
}


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
