// Dense LU with partial pivoting (plain FP64) and the triangular solves, with
// LAPACK xGETRF/xGETRS semantics, plus the block static/moving update of
// solve_linsys.
//
// Reference call sites: src/linsys/mod_linsys.f90:236 (dgetrf, static block),
// 373-465 (dlaswp/dtrsm/dtrsm/dgemm/dgetrf/dlaswp, mixed static/moving),
// 473 (dgetrs, one right-hand side); the pressure system repeats them
// (src/linsys/mod_pressure_equation.f90:251-390).
//
// Structure: right-looking blocked LU.  Each panel is factored by one
// launch of resident blocks that keep their rows of the panel in LDS and
// meet at a grid barrier once per column (pivot search across blocks);
// the row interchanges are then applied to the rest of the matrix, U12 is
// obtained by a unit-lower triangular solve in LDS, and the trailing matrix
// is updated by an FP64 MFMA GEMM.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/dust_gpu_test.h"
#include "common.h"
#include "lu_internal.h"

namespace dust {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Panel factorisation: columns [k, k+jb) of A, rows [k, m).  Block b owns
// rows [k + b R, k + (b+1) R).  Per column: local first-maximum of |a| ->
// candidates published -> barrier -> global pivot (max value, smallest row
// on ties, as idamax) -> swap inside the panel -> scale and rank-1 update.
// ---------------------------------------------------------------------------
struct PanelArgs {
  double *A;
  long lda;
  int m, k, jb, R;
  int *ipiv;      // 1-based, global row numbers
  int *info;
  unsigned *arr;  // [G] arrival words of the grid barrier
  unsigned *err;  // barrier time-out flag
  unsigned gen0;  // barriers passed by earlier panel launches
  double *cval;   // [2][G]
  int *crow;      // [2][G]
  double *crowv;  // [2][G][jb]
  double *topv;   // [2][jb]
};

// (value, row) maximum in idamax order: larger |a|, then the smaller row.
__device__ inline void pivot_max(double &v, int &i, double ov, int oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}

__device__ inline void wave_pivot_max(double &v, int &i) {
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o, 64);
    int oi = __shfl_xor(i, o, 64);
    pivot_max(v, i, ov, oi);
  }
}

__global__ void __launch_bounds__(256) k_panel(PanelArgs P) {
  extern __shared__ double sp[];  // [R][jb] row-major panel rows + pivot rows
  const int G = gridDim.x, b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int jb = P.jb, R = P.R;
  const int r0 = P.k + b * R;
  const int r1 = min(P.m, r0 + R);
  const int nr = max(0, r1 - r0);
  double *rows = sp;               // nr x jb
  double *prow = sp + R * jb;      // jb
  double *trow = prow + jb;        // jb
  __shared__ double wvv[4];
  __shared__ int wvi[4];
  for (int q = tid; q < nr * jb; q += nt) {
    int r = q / jb, c = q % jb;
    rows[q] = P.A[(long)(r0 + r) + P.lda * (P.k + c)];
  }
  __syncthreads();
  for (int jj = 0; jj < jb; ++jj) {
    const int par = jj & 1;
    const int gr = P.k + jj;  // global row of the diagonal
    // block first maximum over rows >= gr: wave butterflies, then 4 partials
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int r = tid; r < nr; r += nt) {
      if (r0 + r < gr) continue;
      pivot_max(best, bi, fabs(rows[r * jb + jj]), r0 + r);
    }
    wave_pivot_max(best, bi);
    if (lane == 0) {
      wvv[wv] = best;
      wvi[wv] = bi;
    }
    __syncthreads();
    best = wvv[0];
    bi = wvi[0];
    for (int w = 1; w < 4; ++w) pivot_max(best, bi, wvv[w], wvi[w]);
    if (tid == 0) {
      P.cval[par * G + b] = best;
      P.crow[par * G + b] = bi;
    }
    if (bi != 0x7fffffff)
      for (int c = tid; c < jb; c += nt) P.crowv[((long)par * G + b) * jb + c] = rows[(bi - r0) * jb + c];
    if (gr >= r0 && gr < r1)
      for (int c = tid; c < jb; c += nt) P.topv[par * jb + c] = rows[(gr - r0) * jb + c];
    if (!grid_sync_arrive(P.arr, P.err, G, P.gen0 + jj + 1)) return;
    // global pivot, computed by every wave from the G candidates
    double gb = -1.0;
    int gi = 0x7fffffff;
    for (int q = lane; q < G; q += 64)
      pivot_max(gb, gi,
                __hip_atomic_load(&P.cval[par * G + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                __hip_atomic_load(&P.crow[par * G + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    wave_pivot_max(gb, gi);
    const int p = gi, pw = (gi - P.k) / R;
    for (int c = tid; c < jb; c += nt) {
      prow[c] = __hip_atomic_load(&P.crowv[((long)par * G + pw) * jb + c], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
      trow[c] = __hip_atomic_load(&P.topv[par * jb + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (p != gr) {
      if (p >= r0 && p < r1)
        for (int c = tid; c < jb; c += nt) rows[(p - r0) * jb + c] = trow[c];
      if (gr >= r0 && gr < r1)
        for (int c = tid; c < jb; c += nt) rows[(gr - r0) * jb + c] = prow[c];
      __syncthreads();
    }
    const double piv = prow[jj];
    if (b == 0 && tid == 0) {
      P.ipiv[gr] = p + 1;
      if (piv == 0.0 && *P.info == 0) *P.info = gr + 1;
    }
    if (piv != 0.0) {
      // dgetf2: scale by the reciprocal when it is representable
      const bool recip = fabs(piv) >= 2.2250738585072014e-308;
      const double rp = 1.0 / piv;
      for (int r = tid; r < nr; r += nt) {
        if (r0 + r <= gr) continue;
        double l = recip ? rows[r * jb + jj] * rp : rows[r * jb + jj] / piv;
        rows[r * jb + jj] = l;
      }
      __syncthreads();
      const int w = jb - jj - 1;
      for (int q = tid; q < nr * w; q += nt) {
        int r = q / w, c = jj + 1 + q % w;
        if (r0 + r <= gr) continue;
        rows[r * jb + c] -= rows[r * jb + jj] * prow[c];
      }
    }
    __syncthreads();
  }
  for (int q = tid; q < nr * jb; q += nt) {
    int r = q / jb, c = q % jb;
    P.A[(long)(r0 + r) + P.lda * (P.k + c)] = rows[q];
  }
}

// Row interchanges ipiv[k..k+jb) applied to the columns outside the panel
// (dlaswp): one thread per column.
__global__ void k_laswp(double *A, long lda, int ncol, int c0, int c1, int k, int jb,
                        const int *__restrict__ ipiv) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol || (c >= c0 && c < c1)) return;
  double *col = A + lda * c;
  for (int i = k; i < k + jb; ++i) {
    int p = ipiv[i] - 1;
    if (p != i) {
      double t = col[i];
      col[i] = col[p];
      col[p] = t;
    }
  }
}

// U12 = L11^-1 A12 for 64 columns per block; L11 unit lower (jb x jb) in LDS.
__global__ void __launch_bounds__(256) k_trsm_llnu(const double *__restrict__ L, long ldl, double *B,
                                                   long ldb, int jb, int ncol) {
  extern __shared__ double st[];
  double *sL = st;              // jb x jb, column-major
  double *sX = st + jb * jb;    // jb x 64, column-major
  const int c0 = blockIdx.x * 64;
  const int nc = min(64, ncol - c0);
  for (int q = threadIdx.x; q < jb * jb; q += blockDim.x) sL[q] = L[(q % jb) + ldl * (q / jb)];
  for (int q = threadIdx.x; q < jb * nc; q += blockDim.x)
    sX[q] = B[(q % jb) + ldb * (long)(c0 + q / jb)];
  __syncthreads();
  for (int i = 0; i < jb - 1; ++i) {
    const int w = jb - 1 - i;
    for (int q = threadIdx.x; q < w * nc; q += blockDim.x) {
      int r = i + 1 + q % w, c = q / w;
      sX[r + jb * c] -= sL[r + jb * i] * sX[i + jb * c];
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < jb * nc; q += blockDim.x)
    B[(q % jb) + ldb * (long)(c0 + q / jb)] = sX[q];
}

// C(M,N) -= A(M,K) B(K,N), FP64 on the matrix cores (v_mfma_f64_16x16x4f64).
// Block tile 64x64, four waves in 2x2, each wave a 32x32 tile of 2x2 MFMA
// tiles; K staged through LDS in chunks of 16.
__global__ void __launch_bounds__(256) k_gemm_f64_sub(int M, int N, int K, const double *__restrict__ A,
                                                      long lda, const double *__restrict__ B, long ldb,
                                                      double *C, long ldc) {
  __shared__ double sA[16][64 + 1];  // [k][m]
  __shared__ double sB[16][64 + 1];  // [k][n]
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv & 1, wc = wv >> 1;
  d4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += 16) {
    // A chunk 64 x 16: thread -> (m = tid % 64, k = tid / 64 + 4 t)
    for (int t = 0; t < 4; ++t) {
      int kk = tid / 64 + 4 * t, mm = tid % 64;
      int gm = m0 + mm, gk = k0 + kk;
      sA[kk][mm] = (gm < M && gk < K) ? A[gm + lda * gk] : 0.0;
    }
    // B chunk 16 x 64: thread -> (k = tid % 16, n = tid / 16 + 16 t)
    for (int t = 0; t < 4; ++t) {
      int kk = tid % 16, nn = tid / 16 + 16 * t;
      int gn = n0 + nn, gk = k0 + kk;
      sB[kk][nn] = (gn < N && gk < K) ? B[gk + ldb * gn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      double af[2], bf[2];
      for (int a = 0; a < 2; ++a) af[a] = sA[kk][wr * 32 + a * 16 + (lane & 15)];
      for (int b = 0; b < 2; ++b) bf[b] = sB[kk][wc * 32 + b * 16 + (lane & 15)];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // FP64 accumulator layout of v_mfma_f64_16x16x4f64: lane holds column
  // (lane % 16) and rows 4 i + lane / 16 (i = register index); the FP32
  // variants use rows 4 (lane / 16) + i instead.
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      int gn = n0 + wc * 32 + b * 16 + (lane & 15);
      if (gn >= N) continue;
      for (int i = 0; i < 4; ++i) {
        int gm = m0 + wr * 32 + a * 16 + 4 * i + (lane >> 4);
        if (gm < M) C[gm + ldc * gn] -= acc[a][b][i];
      }
    }
}

// ---------------------------------------------------------------------------
// getrs: row interchanges on B, then block forward / backward substitution
// with 128-row blocks whose diagonal triangles are inverted first.
// ---------------------------------------------------------------------------
constexpr int SB = 128;   // solve block
constexpr int SR = 8;     // right-hand sides per persistent solve pass

// B rows interchanged by ipiv[0..n) in order (dlaswp, incx = 1).  One block:
// chunks of 256 pivots are screened with ballots so identity entries cost
// nothing; thread q then applies the chunk's interchanges to column q.
__global__ void __launch_bounds__(256) k_laswp_rhs(double *B, long ldb, int nrhs, int n,
                                                   const int *__restrict__ ipiv) {
  __shared__ unsigned long long mask[4];
  __shared__ int pv[256];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int base = 0; base < n; base += 256) {
    const int i = base + tid;
    int p = i < n ? ipiv[i] - 1 : i;
    pv[tid] = p;
    unsigned long long m = __ballot(i < n && p != i);
    if (lane == 0) mask[wv] = m;
    __syncthreads();
    for (int q = tid; q < nrhs; q += 256) {
      double *b = B + ldb * q;
      for (int w = 0; w < 4; ++w) {
        unsigned long long mm = mask[w];
        while (mm) {
          const int t = __builtin_ctzll(mm);
          mm &= mm - 1;
          const int ii = base + 64 * w + t, pp = pv[64 * w + t];
          double x = b[ii];
          b[ii] = b[pp];
          b[pp] = x;
        }
      }
    }
    __syncthreads();
  }
}

// Inverses of the diagonal blocks of the LU factors (dtrti2 column sweeps in
// LDS): Linv = L_kk^-1 (unit lower) and Uinv = U_kk^-1, written row-major.
// Rows/columns past n are padded with the identity.
__global__ void __launch_bounds__(256) k_diag_inverse(const double *__restrict__ A, long lda, int n,
                                                      double *Linv, double *Uinv) {
  extern __shared__ double sd[];  // SB x (SB + 1), column-major
  __shared__ double tl[SB], tu[SB];
  constexpr int LD = SB + 1;
  const int kb = blockIdx.x, k0 = kb * SB, tid = threadIdx.x;
  const int w = min(SB, n - k0);
  for (int q = tid; q < SB * SB; q += 256) {
    int r = q & (SB - 1), c = q >> 7;
    sd[c * LD + r] = (r < w && c < w) ? A[(k0 + r) + lda * (long)(k0 + c)] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  // U: j = 0..SB-1: u_jj <- 1/u_jj; U[0:j, j] <- -u_jj^-1 Uinv[0:j, 0:j] U[0:j, j]
  // L: j = SB-1..0: L[j+1:, j] <- -Linv[j+1:, j+1:] L[j+1:, j]
  for (int s = 0; s < SB; ++s) {
    const int ju = s, jl = SB - 1 - s;
    if (tid < SB) {
      if (tid < ju) tu[tid] = sd[ju * LD + tid];
    } else {
      const int r = tid - SB;
      if (r > jl) tl[r] = sd[jl * LD + r];
    }
    __syncthreads();
    if (tid < SB) {
      const int r = tid;
      const double dj = 1.0 / sd[ju * LD + ju];
      if (r < ju) {
        double acc = 0.0;
        for (int c = r; c < ju; ++c) acc += sd[c * LD + r] * tu[c];
        sd[ju * LD + r] = -dj * acc;
      }
    } else {
      const int r = tid - SB;
      if (r > jl) {
        double acc = tl[r];
        for (int c = jl + 1; c < r; ++c) acc += sd[c * LD + r] * tl[c];
        sd[jl * LD + r] = -acc;
      }
    }
    __syncthreads();
    if (tid == 0) sd[ju * LD + ju] = 1.0 / sd[ju * LD + ju];
    __syncthreads();
  }
  double *ol = Linv + (long)kb * SB * SB, *ou = Uinv + (long)kb * SB * SB;
  for (int q = tid; q < SB * SB; q += 256) {
    int c = q & (SB - 1), r = q >> 7;
    const double v = sd[c * LD + r];
    ol[r * SB + c] = c < r ? v : (c == r ? 1.0 : 0.0);
    ou[r * SB + c] = c >= r ? v : 0.0;
  }
}

// y[i][q] = B[i, q0 + q] (rows past n and columns past nq: 0), and back.
__global__ void k_pack_y(const double *__restrict__ B, long ldb, int n, int np, int nq, double *y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  for (int q = 0; q < SR; ++q) y[(long)i * SR + q] = (i < n && q < nq) ? B[i + ldb * q] : 0.0;
}
__global__ void k_unpack_y(const double *__restrict__ y, int n, int nq, double *B, long ldb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nq; ++q) B[i + ldb * q] = y[(long)i * SR + q];
}

// One word set on the stream (a kernel, not a memset: the solve may be
// recorded into a graph).
__global__ void k_set_word(unsigned *p, unsigned v) {
  if (threadIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------
static int g_cus = 0;
static int device_cus() {
  if (g_cus == 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) == hipSuccess) g_cus = pr.multiProcessorCount;
    if (g_cus <= 0) g_cus = 64;
  }
  return g_cus;
}

int lu_getrf_dev(int m, int n, double *dA, long lda, int *dipiv, hipStream_t st, int nb,
                 int *info_out) {
  DUST_CHECK(m >= 0 && n >= 0 && lda >= std::max(1, m), DUST_E_ARG, "invalid getrf dimensions");
  *info_out = 0;
  if (m == 0 || n == 0) return 0;
  const int cus = device_cus();
  const int Gmax = cus;
  // workspace: barrier + candidates
  size_t wbytes = 256 + sizeof(double) * 2 * Gmax + sizeof(int) * 2 * Gmax +
                  sizeof(double) * 2 * (size_t)Gmax * nb + sizeof(double) * 2 * nb + 64 +
                  sizeof(unsigned) * (Gmax + 16);
  TmpDev<char> wsb;
  DUST_TRY(wsb.alloc(wbytes));
  char *ws = wsb.p;
  DUST_HIP(hipMemsetAsync(ws, 0, wbytes, st));
  unsigned *bar = (unsigned *)ws;
  int *dinfo = (int *)(ws + 16);
  double *cval = (double *)(ws + 256);
  int *crow = (int *)(cval + 2 * Gmax);
  double *crowv = (double *)(crow + 2 * Gmax + 2);
  crowv = (double *)(((uintptr_t)crowv + 15) & ~(uintptr_t)15);
  double *topv = crowv + 2 * (size_t)Gmax * nb;
  unsigned *arr = (unsigned *)(topv + 2 * nb);
  unsigned gen0 = 0;
  const int mn = std::min(m, n);
  int rc = 0;
  for (int k = 0; k < mn && !rc; k += nb) {
    const int jb = std::min(nb, mn - k);
    const int mrows = m - k;
    int G = std::min(std::min(Gmax, 128), std::max(1, (mrows + 63) / 64));
    int R = (mrows + G - 1) / G;
    G = (mrows + R - 1) / R;
    size_t lds = sizeof(double) * ((size_t)R * jb + 2 * jb);
    if (lds > 150 * 1024) {
      set_error("panel of %d x %d does not fit in LDS", R, jb);
      rc = DUST_E_UNSUPPORTED;
      break;
    }
    PanelArgs P{dA, lda, m, k, jb, R, dipiv, dinfo, arr, bar + 2, gen0, cval, crow, crowv, topv};
    gen0 += jb;
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void *)k_panel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
    // G <= 128 blocks of one CU each: all resident at once on an idle GPU (the
    // grid barrier is bounded and reports a time-out otherwise).  An ordinary
    // launch: the kernel does its own synchronisation and needs none of the
    // cooperative-launch resources, which another process may hold.
    k_panel<<<G, 256, lds, st>>>(P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("panel launch failed (G=%d, lds=%zu): %s", G, lds, hipGetErrorString(e));
      rc = DUST_E_HIP;
      break;
    }
    // interchanges on the other columns
    k_laswp<<<ceil_div(n, 256), 256, 0, st>>>(dA, lda, n, k, k + jb, k, jb, dipiv);
    if ((e = hipGetLastError()) != hipSuccess) {
      set_error("laswp launch failed: %s", hipGetErrorString(e));
      rc = DUST_E_HIP;
      break;
    }
    const int nrest = n - k - jb;
    if (nrest > 0) {
      size_t tl = sizeof(double) * ((size_t)jb * jb + (size_t)jb * 64);
      if (tl > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)k_trsm_llnu,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)tl);
      k_trsm_llnu<<<ceil_div(nrest, 64), 256, tl, st>>>(dA + k + lda * k, lda,
                                                        dA + k + lda * (k + jb), lda, jb, nrest);
      const int mrest = m - k - jb;
      if (mrest > 0) {
        dim3 grid(ceil_div(mrest, 64), ceil_div(nrest, 64));
        k_gemm_f64_sub<<<grid, 256, 0, st>>>(mrest, nrest, jb, dA + (k + jb) + lda * k, lda,
                                             dA + k + lda * (k + jb), lda,
                                             dA + (k + jb) + lda * (k + jb), lda);
      }
      if ((e = hipGetLastError()) != hipSuccess) {
        set_error("trailing update launch failed: %s", hipGetErrorString(e));
        rc = DUST_E_HIP;
        break;
      }
    }
  }
  unsigned hb[3] = {0, 0, 0};
  int hinfo = 0;
  if (!rc) {
    DUST_HIP(hipMemcpyAsync(hb, bar, sizeof(hb), hipMemcpyDeviceToHost, st));
    DUST_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, st));
  }
  DUST_HIP(hipStreamSynchronize(st));
  if (rc) return rc;
  DUST_CHECK(hb[2] == 0, DUST_E_HIP, "panel grid barrier timed out");
  *info_out = hinfo;
  return 0;
}

struct GetrsWs {
  DevBuf<double> linv, uinv, y;
  DevBuf<unsigned> flags;  // [0] error, [1..] ready flags (forward, backward)
  unsigned gen = 0;
};
static GetrsWs g_getrs;

void lu_release() {
  g_getrs.linv.release();
  g_getrs.uinv.release();
  g_getrs.y.release();
  g_getrs.flags.release();
  g_getrs.gen = 0;
}

// Test hook: the error word of the getrs workspace (when allocated).
int lu_getrs_set_error_word(unsigned v) {
  if (!g_getrs.flags.p) return 0;
  DUST_HIP(hipMemcpy(g_getrs.flags.p, &v, sizeof(unsigned), hipMemcpyHostToDevice));
  return 0;
}

// Workspace of one getrs: inverted diagonal blocks, the packed right-hand
// sides, the error word and the ready flags, and the generation counter of
// the flags.
struct GetrsPtrs {
  double *linv, *uinv, *y;
  unsigned *flags;  // [0] error, [1..2 nblk] ready flags (forward, backward)
  unsigned *gen;
};

static int getrs_core(int n, int nrhs, const double *dA, long lda, const int *dipiv, double *dB,
                      long ldb, hipStream_t st, const GetrsPtrs &w) {
  const int nblk = ceil_div(n, SB), np = nblk * SB;
  const size_t lds = sizeof(double) * SB * (SB + 1);
  static std::once_flag attr;
  hipError_t ae = hipSuccess;
  std::call_once(attr, [&] {
    ae = hipFuncSetAttribute((const void *)k_diag_inverse, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds);
  });
  DUST_HIP(ae);
  // the error word starts every call at zero: a time-out of an earlier call
  // must not fail this one (the ready flags are generation-stamped)
  k_set_word<<<1, 64, 0, st>>>(w.flags, 0u);
  k_laswp_rhs<<<1, 256, 0, st>>>(dB, ldb, nrhs, n, dipiv);
  k_diag_inverse<<<nblk, 256, lds, st>>>(dA, lda, n, w.linv, w.uinv);
  DUST_HIP(hipGetLastError());
  const int G = std::min(nblk, device_cus());
  for (int q0 = 0; q0 < nrhs; q0 += SR) {
    const int nq = std::min(SR, nrhs - q0);
    k_pack_y<<<ceil_div(np, 256), 256, 0, st>>>(dB + ldb * q0, ldb, n, np, nq, w.y);
    for (int upper = 0; upper < 2; ++upper) {
      if (++*w.gen == 0) ++*w.gen;
      BlockSolveArgs<double> sa{dA, lda, n, nblk, upper ? w.uinv : w.linv, w.y,
                                w.flags + 1 + upper * nblk, *w.gen, w.flags};
      // G <= #CUs small blocks, all resident at once (ordinary launch; the
      // flag waits are bounded)
      if (upper) k_block_solve<double, 1, 512><<<G, 512, 0, st>>>(sa);
      else k_block_solve<double, 0, 512><<<G, 512, 0, st>>>(sa);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        set_error("triangular solve launch failed: %s", hipGetErrorString(e));
        return DUST_E_HIP;
      }
    }
    k_unpack_y<<<ceil_div(n, 256), 256, 0, st>>>(w.y, n, nq, dB + ldb * q0, ldb);
    DUST_HIP(hipGetLastError());
  }
  unsigned herr = 0;
  DUST_HIP(hipMemcpyAsync(&herr, w.flags, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  DUST_HIP(hipStreamSynchronize(st));
  DUST_CHECK(herr == 0, DUST_E_HIP, "triangular solve flag wait timed out");
  return 0;
}

int lu_getrs_dev(int n, int nrhs, const double *dA, long lda, const int *dipiv, double *dB,
                 long ldb, hipStream_t st) {
  DUST_CHECK(n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), DUST_E_ARG,
             "invalid getrs dimensions");
  if (n == 0 || nrhs == 0) return 0;
  GetrsWs &w = g_getrs;
  const int nblk = ceil_div(n, SB), np = nblk * SB;
  DUST_TRY(w.linv.reserve((size_t)nblk * SB * SB));
  DUST_TRY(w.uinv.reserve((size_t)nblk * SB * SB));
  DUST_TRY(w.y.reserve((size_t)np * SR));
  if (w.flags.n < (size_t)(1 + 2 * nblk)) {
    w.flags.n = 0;
    DUST_TRY(w.flags.reserve(1 + 2 * nblk));
    DUST_HIP(hipMemset(w.flags.p, 0, sizeof(unsigned) * w.flags.n));
  }
  return getrs_core(n, nrhs, dA, lda, dipiv, dB, ldb, st,
                    GetrsPtrs{w.linv.p, w.uinv.p, w.y.p, w.flags.p, &w.gen});
}

// ---------------------------------------------------------------------------
// getrs without allocation: the workspace is the file-scope device array
// `scratch` (every GPU holds one copy until the process ends), laid out as
// linv | uinv | y (doubles) | error word + ready flags (unsigned), sized for
// n <= kNoAllocMaxN.  Its address on each device comes from
// hipGetSymbolAddress on the first call with that device current.  Calls on
// one GPU share the array: calls that overlap in time on one GPU race.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxN = 6144;
constexpr int kNoAllocBlk = kNoAllocMaxN / SB;
constexpr size_t kScrInv = (size_t)kNoAllocBlk * SB * SB;           // doubles per inverse set
constexpr size_t kScrY = (size_t)kNoAllocBlk * SB * SR;             // doubles
constexpr size_t kScrDoubles = 2 * kScrInv + kScrY;
constexpr size_t kScrFlags = 1 + 2 * (size_t)kNoAllocBlk;           // unsigned
constexpr size_t kScrBytes = sizeof(double) * kScrDoubles + sizeof(unsigned) * kScrFlags;
static_assert(kNoAllocMaxN % SB == 0, "whole solve blocks");

static __device__ __align__(256) unsigned char scratch[kScrBytes];

constexpr int kMaxDevices = 64;
struct ScratchDev {
  unsigned char *ptr = nullptr;  // device address of `scratch` on this device
  unsigned gen = 0;              // generation of the ready flags
};
static ScratchDev g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_ptrs(GetrsPtrs *out) {
  int dev = 0;
  DUST_HIP(hipGetDevice(&dev));
  DUST_CHECK(dev >= 0 && dev < kMaxDevices, DUST_E_UNSUPPORTED, "device %d beyond %d devices",
             dev, kMaxDevices);
  ScratchDev &s = g_scratch[dev];
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!s.ptr) {
      void *p = nullptr;
      DUST_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      DUST_CHECK(p != nullptr, DUST_E_HIP, "no device address for the getrs scratch array");
      s.ptr = (unsigned char *)p;
    }
  }
  double *d = (double *)s.ptr;
  *out = GetrsPtrs{d, d + kScrInv, d + 2 * kScrInv, (unsigned *)(d + kScrDoubles), &s.gen};
  return 0;
}

int lu_getrs_noalloc_dev(int n, int nrhs, const double *dA, long lda, const int *dipiv,
                         double *dB, long ldb, hipStream_t st) {
  DUST_CHECK(n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), DUST_E_ARG,
             "invalid getrs dimensions");
  DUST_CHECK(n <= kNoAllocMaxN, DUST_E_CAPACITY,
             "dgetrs_noalloc: n = %d exceeds the static workspace (n <= %d)", n, kNoAllocMaxN);
  if (n == 0 || nrhs == 0) return 0;
  GetrsPtrs w;
  DUST_TRY(scratch_ptrs(&w));
  return getrs_core(n, nrhs, dA, lda, dipiv, dB, ldb, st, w);
}

int lu_gemm_sub_dev(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                    double *C, long ldc, hipStream_t st) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  dim3 grid(ceil_div(M, 64), ceil_div(N, 64));
  k_gemm_f64_sub<<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
  DUST_HIP(hipGetLastError());
  return 0;
}

int lu_laswp_dev(double *A, long lda, int ncol, int k, int jb, const int *dipiv, hipStream_t st) {
  if (ncol <= 0 || jb <= 0) return 0;
  k_laswp<<<ceil_div(ncol, 256), 256, 0, st>>>(A, lda, ncol, 0, 0, k, jb, dipiv);
  DUST_HIP(hipGetLastError());
  return 0;
}

int lu_trsm_llnu_dev(int m, int n, const double *L, long ldl, double *B, long ldb, hipStream_t st) {
  // blocked: diagonal blocks of 64 solved in LDS, off-diagonal updates by GEMM
  for (int k = 0; k < m; k += 64) {
    int jb = std::min(64, m - k);
    size_t tl = sizeof(double) * ((size_t)jb * jb + (size_t)jb * 64);
    k_trsm_llnu<<<ceil_div(n, 64), 256, tl, st>>>(L + k + ldl * k, ldl, B + k, ldb, jb, n);
    if (k + jb < m) DUST_TRY(lu_gemm_sub_dev(m - k - jb, n, jb, L + (k + jb) + ldl * k, ldl, B + k, ldb,
                                             B + k + jb, ldb, st));
  }
  DUST_HIP(hipGetLastError());
  return 0;
}

}  // namespace dust

using namespace dust;

extern "C" {

int dust_dgetrf_dev(int m, int n, double *dA, int lda, int *dipiv, void *stream) {
  int info = 0;
  DUST_TRY(lu_getrf_dev(m, n, dA, lda, dipiv, (hipStream_t)stream, 64, &info));
  return info;
}

int dust_dgetrs_dev(int n, int nrhs, const double *dA, int lda, const int *dipiv, double *dB,
                    int ldb, void *stream) {
  DUST_TRY(lu_getrs_dev(n, nrhs, dA, lda, dipiv, dB, ldb, (hipStream_t)stream));
  DUST_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int dust_dgetrf(int m, int n, double *A, int lda, int *ipiv) {
  DUST_CHECK(A && ipiv && m >= 0 && n >= 0 && lda >= std::max(1, m), DUST_E_ARG,
             "invalid dgetrf arguments");
  if (m == 0 || n == 0) return 0;
  const int mn = std::min(m, n);
  TmpDev<double> dA;
  TmpDev<int> dp;
  DUST_TRY(dA.alloc((size_t)m * n));
  DUST_TRY(dp.alloc(mn));
  DUST_HIP(hipMemcpy2D(dA.p, sizeof(double) * m, A, sizeof(double) * lda, sizeof(double) * m, n,
                       hipMemcpyHostToDevice));
  int info = 0;
  DUST_TRY(lu_getrf_dev(m, n, dA.p, m, dp.p, nullptr, 64, &info));
  DUST_HIP(hipMemcpy2D(A, sizeof(double) * lda, dA.p, sizeof(double) * m, sizeof(double) * m, n,
                       hipMemcpyDeviceToHost));
  DUST_HIP(hipMemcpy(ipiv, dp.p, sizeof(int) * mn, hipMemcpyDeviceToHost));
  return info;
}

// Test hooks on device pointers: the trailing-update GEMM, the unit-lower
// triangular solve and the row interchanges used inside dgetrf.
int dust_test_gemm_sub_dev(int M, int N, int K, const double *A, int lda, const double *B, int ldb,
                           double *C, int ldc) {
  DUST_TRY(lu_gemm_sub_dev(M, N, K, A, lda, B, ldb, C, ldc, nullptr));
  DUST_HIP(hipDeviceSynchronize());
  return 0;
}
int dust_test_trsm_llnu_dev(int m, int n, const double *L, int ldl, double *B, int ldb) {
  DUST_TRY(lu_trsm_llnu_dev(m, n, L, ldl, B, ldb, nullptr));
  DUST_HIP(hipDeviceSynchronize());
  return 0;
}
int dust_test_getrs_set_error_word(unsigned v) { return lu_getrs_set_error_word(v); }

int dust_dgetrs_noalloc_dev(int n, int nrhs, const double *A, int lda, const int *ipiv,
                            double *B, int ldb, void *stream) {
  DUST_CHECK(A && ipiv && B, DUST_E_ARG, "null dgetrs_noalloc argument");
  return lu_getrs_noalloc_dev(n, nrhs, A, lda, ipiv, B, ldb, (hipStream_t)stream);
}

int dust_dgetrs(int n, int nrhs, const double *A, int lda, const int *ipiv, double *B, int ldb) {
  DUST_CHECK(A && ipiv && B && n >= 0 && nrhs >= 0 && lda >= std::max(1, n) &&
                 ldb >= std::max(1, n),
             DUST_E_ARG, "invalid dgetrs arguments");
  if (n == 0 || nrhs == 0) return 0;
  TmpDev<double> dA, dB;
  TmpDev<int> dp;
  DUST_TRY(dA.alloc((size_t)n * n));
  DUST_TRY(dB.alloc((size_t)n * nrhs));
  DUST_TRY(dp.alloc(n));
  DUST_HIP(hipMemcpy2D(dA.p, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                       hipMemcpyHostToDevice));
  DUST_HIP(hipMemcpy2D(dB.p, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                       hipMemcpyHostToDevice));
  DUST_HIP(hipMemcpy(dp.p, ipiv, sizeof(int) * n, hipMemcpyHostToDevice));
  DUST_TRY(lu_getrs_dev(n, nrhs, dA.p, n, dp.p, dB.p, n, nullptr));
  DUST_HIP(hipMemcpy2D(B, sizeof(double) * ldb, dB.p, sizeof(double) * n, sizeof(double) * n, nrhs,
                       hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
