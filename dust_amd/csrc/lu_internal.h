// Device-pointer LU building blocks shared by the LU translation units.
#pragma once
#include "common.h"

namespace dust {

// Bounded spin until *f == gen (acquire).  On time-out raises *err; returns
// false when *err is raised by anyone.
__device__ inline bool wait_flag(unsigned *f, unsigned gen, unsigned *err) {
  long spins = 0;
  while (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != gen) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins > (1L << 25)) {
      atomicExch(err, 1u);
      return false;
    }
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
  }
  return true;
}
// Grid barrier for a grid whose blocks are all resident at once (a grid of
// at most one block per CU), without a shared counter: block b publishes
// arrival number
// `target` in arr[b] and wave 0 polls all G words (G <= 128).  A block that
// has already left this barrier can only have published target + 1, so the
// test is "arr[q] - target >= 0" in wrap-around arithmetic.  Bounded spin;
// a time-out raises *err.
__device__ inline bool grid_sync_arrive(unsigned *arr, unsigned *err, int G, unsigned target) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_store(&arr[blockIdx.x], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    long spins = 0;
    int good = 1;
    while (true) {
      bool mine = true;
      for (int q = lane; q < G; q += 64)
        mine &= (int)(__hip_atomic_load(&arr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) >= 0;
      if (__all(mine)) break;
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1L << 26)) {
        if (lane == 0) atomicExch(err, 1u);
        good = 0;
        break;
      }
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        good = 0;
        break;
      }
    }
    if (lane == 0) {
      __threadfence();
      ok = good;
    }
  }
  __syncthreads();
  return ok != 0;
}

// Workgroup `bid` of a batched launch belongs to the item i with
// starts[i] <= bid < starts[i + 1] (starts ascending, starts[0] = 0; items
// own at least one workgroup).
__device__ inline int batch_item(const int *__restrict__ starts, int m, int bid) {
  int lo = 0, hi = m;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (starts[mid] <= bid) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Block substitution with 128-row blocks for up to 8 right-hand sides held as
// y[row][8], one workgroup per block row in dependency order (forward: L unit
// lower from the off-diagonal blocks of A; backward: U).  Every block row
// waits on the ready flag (generation-stamped) of each block row it reads.
// Each thread owns CW columns of its row: the row segment of every
// off-diagonal block and of the diagonal-block inverse is loaded before the
// wait, so a ready flag is followed only by the 4 KB y block and FMAs.
// Inv holds the inverse of every diagonal block, row-major (a thread's row
// segment is contiguous).
template <typename T>
struct BlockSolveArgs {
  const T *A;
  long lda;
  int n, nblk;
  const T *Inv;       // [nblk][128*128], row-major blocks
  T *y;               // [nblk*128][8], in place
  unsigned *ready;    // [nblk]
  unsigned gen;
  unsigned *err;
};

// Block rows bid, bid + G, ... of one system (G workgroups share it).
template <typename T, int UPPER, int NT>
__device__ __forceinline__ void block_solve_rows(const BlockSolveArgs<T> &a, int bid, int G) {
  constexpr int BS = 128, NQ = NT / BS, CW = BS / NQ, R8 = 8;
  __shared__ T sy[BS * R8];
  __shared__ T sp[NQ * BS * R8];
  __shared__ int ok;
  const int t = threadIdx.x, r = t & (BS - 1), h = t / BS;
  for (int it = bid; it < a.nblk; it += G) {
    const int kb = UPPER ? a.nblk - 1 - it : it;
    const int gr = kb * BS + r;
    const bool rok = gr < a.n;
    T inv[CW];
    {
      const T *Iv = a.Inv + (long)kb * BS * BS + r * BS + h * CW;
#pragma unroll
      for (int c = 0; c < CW; ++c) inv[c] = Iv[c];
    }
    T acc[R8];
#pragma unroll
    for (int q = 0; q < R8; ++q) acc[q] = T(0);
    const int nj = UPPER ? a.nblk - 1 - kb : kb;
    for (int jj = 0; jj < nj; ++jj) {
      const int j = UPPER ? a.nblk - 1 - jj : jj;
      T l[CW];
      const int c0 = j * BS + h * CW;
      const int ldi = (int)a.lda;  // column offsets inside one block fit in 32 bits
      // clamped addresses, no branches between the loads; columns past n
      // (the last block) are zeroed afterwards
      const T *Lc = a.A + (rok ? gr : 0) + a.lda * (long)min(c0, a.n - 1);
      const int cmax = a.n - 1 - min(c0, a.n - 1);
#pragma unroll
      for (int c = 0; c < CW; ++c) l[c] = Lc[ldi * min(c, cmax)];
#pragma unroll
      for (int c = 0; c < CW; ++c)
        if (!rok || c0 + c >= a.n) l[c] = T(0);
      if (t == 0) ok = wait_flag(&a.ready[j], a.gen, a.err) ? 1 : 0;
      __syncthreads();
      if (!ok) return;
      {
        T yv[BS * R8 / NT];
#pragma unroll
        for (int u = 0; u < BS * R8 / NT; ++u) yv[u] = a.y[(long)j * BS * R8 + t + NT * u];
#pragma unroll
        for (int u = 0; u < BS * R8 / NT; ++u) sy[t + NT * u] = yv[u];
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        if ((c & 3) == 0) __builtin_amdgcn_sched_barrier(0);  // keep the LDS reads near their use
#pragma unroll
        for (int q = 0; q < R8; ++q) acc[q] += l[c] * sy[(h * CW + c) * R8 + q];
      }
      __syncthreads();
    }
    // right-hand side minus the partial sums of the NQ column groups
#pragma unroll
    for (int q = 0; q < R8; ++q) sp[(h * BS + r) * R8 + q] = acc[q];
    __syncthreads();
    if (h == 0) {
#pragma unroll
      for (int q = 0; q < R8; ++q) {
        T v = a.y[(long)gr * R8 + q];
        for (int g = 0; g < NQ; ++g) v -= sp[(g * BS + r) * R8 + q];
        sy[r * R8 + q] = v;
      }
    }
    __syncthreads();
    // diagonal block: y_kb = Inv_kb (rhs)
#pragma unroll
    for (int q = 0; q < R8; ++q) acc[q] = T(0);
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      if ((c & 3) == 0) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < R8; ++q) acc[q] += inv[c] * sy[(h * CW + c) * R8 + q];
    }
#pragma unroll
    for (int q = 0; q < R8; ++q) sp[(h * BS + r) * R8 + q] = acc[q];
    __syncthreads();
    if (h == 0) {
#pragma unroll
      for (int q = 0; q < R8; ++q) {
        T v = sp[r * R8 + q];
        for (int g = 1; g < NQ; ++g) v += sp[(g * BS + r) * R8 + q];
        a.y[(long)gr * R8 + q] = v;
      }
      __threadfence();
    }
    __syncthreads();
    if (t == 0) __hip_atomic_store(&a.ready[kb], a.gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename T, int UPPER, int NT>
__global__ void __launch_bounds__(NT) k_block_solve(BlockSolveArgs<T> a) {
  block_solve_rows<T, UPPER, NT>(a, blockIdx.x, gridDim.x);
}

// Independent systems in one launch: items[i] owns workgroups
// [starts[i], starts[i+1]).  A workgroup waits only on workgroups of its own
// system; with workgroups dispatched in index order, every system whose
// first workgroup is resident has all its workgroups resident once the
// systems before it have drained, so the waits end.
template <typename T, int UPPER, int NT>
__global__ void __launch_bounds__(NT) k_block_solve_batch(const BlockSolveArgs<T> *__restrict__ items,
                                                          const int *__restrict__ starts, int m) {
  const int i = batch_item(starts, m, blockIdx.x);
  const BlockSolveArgs<T> a = items[i];
  block_solve_rows<T, UPPER, NT>(a, blockIdx.x - starts[i], starts[i + 1] - starts[i]);
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
