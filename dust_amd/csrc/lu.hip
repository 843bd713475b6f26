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
#include <string>
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
  const int *spec_fail;  // NULL, or the speculative factorisation's verdict
                         // (0: it stands and this launch does nothing)
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
  if (P.spec_fail && *P.spec_fail == 0) return;  // uniform over the grid
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

// ---------------------------------------------------------------------------
// Speculative panel: the 64 x 64 diagonal block factored with interchanges
// among its own rows, then L21 = A21 U11^-1 for all rows below.  Partial
// pivoting over the whole panel would have chosen the same pivots exactly
// when no row below ever offers a strictly larger candidate, |a_ij^(j)| >
// |u_jj| (idamax takes the first maximum, and the block rows come first), so
// then the result stands.  Both kernels work in place and keep the original
// values in a backup; when the speculation fails (or meets a pivot below the
// smallest normal number, an infinity or a NaN) the host restores the panel
// and factors it again with k_panel.
// ---------------------------------------------------------------------------
constexpr int SP = 64;  // speculative block rows = panel width
constexpr double kSfmin = 2.2250738585072014e-308;  // dlamch('S')

typedef unsigned long long u64;

// Cross-lane moves inside rows of 16 lanes (DPP): quad_perm [1,0,3,2] and
// [2,3,0,1], row_half_mirror, row_mirror.  All four are permutations, so
// every lane receives a value.
template <int CTRL>
__device__ inline u64 dpp64(u64 v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, 0xf, 0xf, false);
  return ((u64)(unsigned)hi << 32) | (unsigned)lo;
}

__device__ inline u64 readlane64(u64 v, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), l);
  return ((u64)hi << 32) | lo;
}

__device__ inline double readlane_d(double v, int l) {
  return __longlong_as_double((long long)readlane64((u64)__double_as_longlong(v), l));
}

// Maximum (MAX) or minimum over the wavefront, the same value in every lane.
template <bool MAX>
__device__ inline u64 wave_ext_u64(u64 v) {
  auto pick = [](u64 a, u64 b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
  v = pick(v, dpp64<0xB1>(v));
  v = pick(v, dpp64<0x4E>(v));
  v = pick(v, dpp64<0x141>(v));
  v = pick(v, dpp64<0x140>(v));
  return pick(pick(readlane64(v, 0), readlane64(v, 16)), pick(readlane64(v, 32), readlane64(v, 48)));
}

// X = L^-1 for the unit lower triangle of sD (rows and columns < 64; the
// diagonal and upper part are not read), by 256 threads: lane c is column c,
// wave q rows [16q, 16q+16), one 16-row block substitution per wave and the
// updates from earlier blocks through sX.  X(64 x 64) column-major.
__device__ inline void unit_lower_inverse(const double (*sD)[SP + 1], double (*sX)[SP + 1],
                                          double *X) {
  const int q = threadIdx.x >> 6, c = threadIdx.x & 63;
  double x[16];
#pragma unroll
  for (int ii = 0; ii < 16; ++ii) x[ii] = (16 * q + ii == c) ? 1.0 : 0.0;
  for (int p = 0; p < 4; ++p) {
    if (q == p) {
#pragma unroll
      for (int kk = 0; kk < 15; ++kk)
#pragma unroll
        for (int ii = kk + 1; ii < 16; ++ii) x[ii] -= sD[16 * p + ii][16 * p + kk] * x[kk];
#pragma unroll
      for (int ii = 0; ii < 16; ++ii) sX[16 * p + ii][c] = x[ii];
    }
    __syncthreads();
    if (q > p) {
#pragma unroll 2
      for (int kk = 0; kk < 16; ++kk) {
        const double y = sX[16 * p + kk][c];
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) x[ii] -= sD[16 * q + ii][16 * p + kk] * y;
      }
    }
  }
#pragma unroll
  for (int ii = 0; ii < 16; ++ii) X[(16 * q + ii) + SP * c] = x[ii];
}

// Factor the 64 x 64 diagonal block at (k, k) in place with interchanges
// among its rows (dgetf2 order: per column, the first maximum of |a| among
// the rows not yet pivoted, scaling by the reciprocal pivot, rank-1 update).
// 256 threads: wave w holds columns [16w, 16w+16) of the block, lane r row r,
// in registers; rows never move, every lane carries the current position of
// its row (an interchange swaps two labels).  In phase p wave p factors its
// 16 columns alone (DPP maximum of |a| as 64-bit keys, ties to the smallest
// position; the pivot row from its lane by readlane), then publishes the
// pivot lanes, positions and multipliers through LDS; after the barrier the
// waves to its right replay the 16 eliminations on their columns and the
// waves to its left follow the interchanges: four barriers in all.  Writes
// the original block to Dorig, the factored block to A, the interchanges to
// ipiv (1-based, global rows) and L^-1 to Linv; a bad pivot sets *fail and
// leaves A unchanged.
__global__ void __launch_bounds__(256) k_spec_diag(double *A, long lda, int k, double *Dorig,
                                                   int *ipiv, double *Linv, int *fail) {
  __shared__ double sl[2][SP][17];  // multipliers of a phase, [row][column in phase]
  __shared__ int sbl[SP], sbp[SP];  // pivot lane, pivot position per column
  __shared__ int sbad;
  __shared__ double sD[SP][SP + 1];
  __shared__ double sX[SP][SP + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x == 0) sbad = 0;
  double a[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) a[c] = A[(long)(k + lane) + lda * (k + 16 * w + c)];
#pragma unroll
  for (int c = 0; c < 16; ++c) Dorig[lane + SP * (16 * w + c)] = a[c];
  int pos = lane;
  for (int ph = 0; ph < 4; ++ph) {
    if (w == ph) {
      bool bad = false;
#pragma unroll
      for (int jl = 0; jl < 16; ++jl) {
        const int j = 16 * ph + jl;
        const bool elig = pos >= j;
        const double v = a[jl];
        const u64 key = elig ? (u64)__double_as_longlong(fabs(v)) : 0ull;
        const u64 kmax = wave_ext_u64<true>(key);
        const u64 m = __ballot(elig && key == kmax);
        int bl;
        if (__popcll(m) == 1) {
          bl = __ffsll((long long)m) - 1;
        } else {  // tie: the smallest position among the maxima
          const bool in = (m >> lane) & 1ull;
          const u64 pmin = wave_ext_u64<false>(in ? (u64)pos : ~0ull);
          bl = __ffsll((long long)__ballot(in && (u64)pos == pmin)) - 1;
        }
        bl = __builtin_amdgcn_readfirstlane(bl);
        const double piv = readlane_d(v, bl);
        const int bp = __builtin_amdgcn_readlane(pos, bl);
        bad |= kmax == 0ull || kmax >= 0x7ff0000000000000ull || !(fabs(piv) >= kSfmin);
        double prow[16];
#pragma unroll
        for (int c = jl + 1; c < 16; ++c) prow[c] = readlane_d(a[c], bl);
        if (pos == j)
          pos = bp;
        if (lane == bl) pos = j;
        if (pos > j) {
          const double l = v * (1.0 / piv);
          a[jl] = l;
#pragma unroll
          for (int c = jl + 1; c < 16; ++c) a[c] -= l * prow[c];
        }
        if (lane == 0) {
          sbl[j] = bl;
          sbp[j] = bp;
          ipiv[k + j] = k + bp + 1;
        }
      }
#pragma unroll
      for (int jl = 0; jl < 16; ++jl) sl[ph & 1][lane][jl] = a[jl];
      if (bad && lane == 0) sbad = 1;
    }
    __syncthreads();
    if (sbad) {
      if (threadIdx.x == 0) atomicOr(fail, 1);
      return;  // uniform over the block
    }
    if (w != ph) {
      for (int jl = 0; jl < 16; ++jl) {
        const int j = 16 * ph + jl;
        const int bl = __builtin_amdgcn_readfirstlane(sbl[j]);
        const int bp = sbp[j];
        if (w > ph) {
          double prow[16];
#pragma unroll
          for (int c = 0; c < 16; ++c) prow[c] = readlane_d(a[c], bl);
          if (pos == j)
            pos = bp;
          if (lane == bl) pos = j;
          if (pos > j) {
            const double l = sl[ph & 1][lane][jl];
#pragma unroll
            for (int c = 0; c < 16; ++c) a[c] -= l * prow[c];
          }
        } else {
          if (pos == j)
            pos = bp;
          if (lane == bl) pos = j;
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    A[(long)(k + pos) + lda * (k + 16 * w + c)] = a[c];
    sD[pos][16 * w + c] = a[c];
  }
  __syncthreads();
  unit_lower_inverse(sD, sX, Linv);
}

// L^-1 of the unit lower triangle of the jb x jb block at A (jb <= 64; the
// rest of the 64 x 64 result is the identity), for a panel that k_panel
// factored.
__global__ void __launch_bounds__(256) k_trtri_unit(const double *__restrict__ A, long lda, int jb,
                                                    double *Linv) {
  __shared__ double sD[SP][SP + 1];
  __shared__ double sX[SP][SP + 1];
  {
    // rows and columns clamped to the block (loads without branches), then
    // the strict lower triangle kept
    const int r = threadIdx.x & 63, rr = min(r, jb - 1);
    double t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = A[rr + lda * min((int)(threadIdx.x >> 6) + 4 * i, jb - 1)];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int c = (int)(threadIdx.x >> 6) + 4 * i;
      sD[r][c] = (r < jb && c < r) ? t[i] : 0.0;
    }
  }
  __syncthreads();
  unit_lower_inverse(sD, sX, Linv);
}

// L21 = A21 U11^-1 in place for the nrow rows below the factored diagonal
// block, one thread per row (dgetf2 order: scale by the reciprocal pivot,
// then subtract from the later columns), the original rows into W(ldw, 64).
// A row whose candidate |a_ij^(j)| exceeds |u_jj| fails the speculation.
__global__ void __launch_bounds__(64) k_spec_l21(double *__restrict__ A, long lda, int k, int nrow,
                                                 double *__restrict__ W, long ldw, int *fail) {
  __shared__ double su[SP][SP + 1];
  __shared__ double sr[SP];
  for (int c0 = 0; c0 < SP; c0 += 16) {
    double t[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) t[c] = A[(long)(k + threadIdx.x) + lda * (k + c0 + c)];
#pragma unroll
    for (int c = 0; c < 16; ++c) su[threadIdx.x][c0 + c] = (int)threadIdx.x <= c0 + c ? t[c] : 0.0;
  }
  __syncthreads();
  sr[threadIdx.x] = 1.0 / su[threadIdx.x][threadIdx.x];
  __syncthreads();
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= nrow) return;
  const long row = k + SP + r;
  double a[SP];
#pragma unroll
  for (int j = 0; j < SP; ++j) a[j] = A[row + lda * (k + j)];
#pragma unroll
  for (int j = 0; j < SP; ++j) W[r + ldw * j] = a[j];
  bool bad = false;
#pragma unroll
  for (int i = 0; i < SP; ++i) {
    bad |= !(fabs(a[i]) <= fabs(su[i][i]));
    a[i] *= sr[i];
#pragma unroll
    for (int j = i + 1; j < SP; ++j) a[j] -= a[i] * su[i][j];
  }
#pragma unroll
  for (int j = 0; j < SP; ++j) A[row + lda * (k + j)] = a[j];
  if (bad) atomicOr(fail, 1);
}

// The speculation failed: the panel's original columns back into A.
__global__ void k_spec_restore(double *A, long lda, int k, int nrow, const double *__restrict__ Dorig,
                               const double *__restrict__ W, long ldw) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nd = (long)SP * SP, nl = (long)nrow * SP;
  if (q < nd) {
    const int r = (int)(q % SP), c = (int)(q / SP);
    A[(long)(k + r) + lda * (k + c)] = Dorig[r + SP * c];
  } else if (q < nd + nl) {
    const long t = q - nd;
    const int r = (int)(t % nrow), c = (int)(t / nrow);
    A[(long)(k + SP + r) + lda * (k + c)] = W[r + ldw * c];
  }
}

// B(jb, ncol) <- L^-1 B with X = L^-1 (64 x 64 column-major, from
// k_spec_diag or k_trtri_unit), jb <= 64: 256 threads per 64 columns, the
// column tile staged in LDS, lane c computing column c's rows [16q, 16q+16)
// for wave q.
__global__ void __launch_bounds__(256) k_trsm_inv(const double *__restrict__ X, int jb,
                                                  double *__restrict__ B, long ldb, int ncol) {
  __shared__ double sX[SP][SP + 1];
  __shared__ double sB[SP][SP + 1];
  const int c0 = blockIdx.x * 64;
  {
    // rows and columns clamped to the tile (loads without branches)
    const int r = threadIdx.x & 63, rr = min(r, jb - 1);
    double tx[16], tb[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int c = (int)(threadIdx.x >> 6) + 4 * i;
      tx[i] = X[r + SP * c];
      tb[i] = B[rr + ldb * (long)min(c0 + c, ncol - 1)];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int c = (int)(threadIdx.x >> 6) + 4 * i;
      sX[r][c] = tx[i];
      sB[r][c] = (r < jb && c0 + c < ncol) ? tb[i] : 0.0;
    }
  }
  __syncthreads();
  const int q = threadIdx.x >> 6, c = threadIdx.x & 63;
  double y[16];
#pragma unroll
  for (int ii = 0; ii < 16; ++ii) y[ii] = 0.0;
  const int kend = min(16 * q + 16, jb);
  for (int kk = 0; kk < kend; ++kk) {
    const double b = sB[kk][c];
#pragma unroll
    for (int ii = 0; ii < 16; ++ii) y[ii] += sX[16 * q + ii][kk] * b;
  }
  if (c0 + c >= ncol) return;
#pragma unroll
  for (int ii = 0; ii < 16; ++ii)
    if (16 * q + ii < jb) B[(16 * q + ii) + ldb * (long)(c0 + c)] = y[ii];
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

// B(jb, ncol) <- L^-1 B with L unit lower (jb <= 64): one thread per column,
// its column in registers shifted by one row per step (the finished entry is
// always x[0], constant register indices), L in LDS (every lane reads the
// same element: broadcast); no barrier inside the substitution.
__global__ void __launch_bounds__(64) k_trsm_llnu(const double *__restrict__ L, long ldl, double *B,
                                                  long ldb, int jb, int ncol) {
  __shared__ double sL[2 * 64][65];  // rows past jb are zero
  for (int q = threadIdx.x; q < 2 * 64 * 64; q += 64) {
    const int r = q & 127, c = q >> 7;
    sL[r][c] = (r < jb && c < jb) ? L[r + ldl * c] : 0.0;
  }
  __syncthreads();
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= ncol) return;
  double *b = B + ldb * (long)col;
  double x[64];
#pragma unroll
  for (int i = 0; i < 64; ++i) x[i] = i < jb ? b[i] : 0.0;
  for (int i = 0; i < jb; ++i) {
    // x[0] is the finished entry i; eliminate it from the rows below
#pragma unroll
    for (int r = 1; r < 64; ++r) x[r] -= sL[i + r][i] * x[0];
    b[i] = x[0];
#pragma unroll
    for (int r = 0; r < 63; ++r) x[r] = x[r + 1];
    x[63] = 0.0;
  }
}

// C(M,N) -= A(M,K) B(K,N), column-major FP64, on the matrix cores
// (v_mfma_f64_16x16x4f64).  Block tile 128 x 128, four waves in 2 x 2, each a
// 64 x 64 tile of 4 x 4 MFMA tiles; K in chunks of 16 staged through LDS in
// two buffers, so the next chunk's global loads are in flight while the
// current chunk's MFMAs run (one barrier per chunk).  VEC: 16-byte loads
// (even M, K and leading dimensions, 16-byte aligned A and B).
constexpr int GT = 128, GK = 16, GPAD = 4;
typedef double dbl2 __attribute__((ext_vector_type(2)));

template <int VEC>
__global__ void __launch_bounds__(256, 2) k_dgemm_sub(int M, int N, int K, const double *__restrict__ A,
                                                   long lda, const double *__restrict__ B, long ldb,
                                                   double *C, long ldc) {
  __shared__ double sA[2][GK][GT + GPAD];  // [k][m]
  __shared__ double sB[2][GK][GT + GPAD];  // [k][n]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv & 1, wc = wv >> 1;
  const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
  // staging: A chunk rows ar, ar + 1 of columns k0 + ak + 4 q; B chunk rows
  // k0 + bk, bk + 1 of columns bn + 32 q
  const int ar = 2 * (tid & 63), ak = tid >> 6;
  const int bk = 2 * (tid & 7), bn = tid >> 3;
  dbl2 ra[4], rb[4];
  // addresses clamped into the matrices, out-of-range values replaced by
  // zeros only when they are stored to LDS: no branches or uses next to the
  // loads, so the next chunk's loads stay in flight while the current
  // chunk's MFMAs run
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kc = min(k0 + ak + 4 * q, K - 1), r = m0 + ar;
      const int kb = k0 + bk, cc = min(n0 + bn + 32 * q, N - 1);
      if (VEC) {
        // M and K even: a pair starting at an even index is in range or not
        ra[q] = *(const dbl2 *)&A[min(r, M - 2) + lda * kc];
        rb[q] = *(const dbl2 *)&B[min(kb, K - 2) + ldb * cc];
      } else {
        ra[q].x = A[min(r, M - 1) + lda * kc];
        ra[q].y = A[min(r + 1, M - 1) + lda * kc];
        rb[q].x = B[min(kb, K - 1) + ldb * cc];
        rb[q].y = B[min(kb + 1, K - 1) + ldb * cc];
      }
    }
  };
  auto sstore = [&](int buf, int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + ak + 4 * q, r = m0 + ar;
      const int kb = k0 + bk, c = n0 + bn + 32 * q;
      sA[buf][ak + 4 * q][ar] = (k < K && r < M) ? ra[q].x : 0.0;
      sA[buf][ak + 4 * q][ar + 1] = (k < K && r + 1 < M) ? ra[q].y : 0.0;
      sB[buf][bk][bn + 32 * q] = (kb < K && c < N) ? rb[q].x : 0.0;
      sB[buf][bk + 1][bn + 32 * q] = (kb + 1 < K && c < N) ? rb[q].y : 0.0;
    }
  };
  const int nch = (K + GK - 1) / GK;
  gload(0);
  sstore(0, 0);
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    const int buf = c & 1;
    if (c + 1 < nch) gload((c + 1) * GK);
#pragma unroll
    for (int ks = 0; ks < GK / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      double af[4], bf[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) af[a] = sA[buf][kk][wr * 64 + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 4; ++b) bf[b] = sB[buf][kk][wc * 64 + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    if (c + 1 < nch) sstore(buf ^ 1, (c + 1) * GK);
    __syncthreads();
  }
  // FP64 accumulator layout of v_mfma_f64_16x16x4f64: lane holds column
  // (lane % 16) and rows 4 i + lane / 16 (i = register index).  C in four
  // batches of 16 values per lane: all loads of a batch (clamped addresses),
  // then its stores.
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double cv[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int gn = min(n0 + wc * 64 + b * 16 + (lane & 15), N - 1);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gm = min(m0 + wr * 64 + a * 16 + 4 * i + (lane >> 4), M - 1);
        cv[b][i] = C[gm + ldc * gn];
      }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int gn = n0 + wc * 64 + b * 16 + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gm = m0 + wr * 64 + a * 16 + 4 * i + (lane >> 4);
        if (gm < M && gn < N) C[gm + ldc * gn] = cv[b][i] - acc[a][b][i];
      }
    }
  }
}

static void gemm_sub(int M, int N, int K, const double *A, long lda, const double *B, long ldb,
                     double *C, long ldc, hipStream_t st) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  dim3 grid(ceil_div(M, GT), ceil_div(N, GT));
  const bool vec = (M % 2 == 0) && (K % 2 == 0) && (lda % 2 == 0) && (ldb % 2 == 0) &&
                   (((uintptr_t)A & 15) == 0) && (((uintptr_t)B & 15) == 0);
  if (vec) k_dgemm_sub<1><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
  else k_dgemm_sub<0><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
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

// Blocked right-looking LU with LAPACK's pivots: outer panels of NB = 256
// columns factored as 64-column sub-panels (speculative first; k_panel when
// the speculation fails), the sub-panels' updates inside the outer panel,
// then a trailing update of K = 256 per outer panel (a quarter of the
// trailing-matrix traffic of 64-column panels).  One panel of lookahead:
// the next outer panel's columns are updated first and that panel is factored
// on stream s1 while the rest of the trailing matrix is updated on stream s2,
// which is kept off the first g_lu_reserved_cus compute units (a CU-masked
// stream) so that the single-workgroup panel kernels always find a CU.  The
// speculation's verdict is read back per sub-panel (s1 only; s2 keeps
// running); a failed one waits for s2 and refactors the sub-panel with
// k_panel, whose grid barrier needs all its blocks resident at once.
constexpr int LU_NB = 256;
int g_lu_reserved_cus = 32;  // tuning knob "getrf_reserved_cus"

struct GetrfWs {
  int dev = -1;
  hipStream_t s1 = nullptr, s2 = nullptr;
  hipEvent_t ev_in = nullptr, ev_u12 = nullptr, ev_rest = nullptr, ev_done = nullptr;
  HostBuf<int> hflag;
};
static GetrfWs g_getrf;

static int getrf_streams(GetrfWs &w) {
  if (w.s1) return 0;
  DUST_HIP(hipStreamCreateWithFlags(&w.s1, hipStreamNonBlocking));
  const int cus = device_cus();
  std::vector<uint32_t> mask((cus + 31) / 32, 0u);
  for (int c = g_lu_reserved_cus; c < cus; ++c) mask[c / 32] |= 1u << (c % 32);
  if (g_lu_reserved_cus <= 0 || cus <= 2 * g_lu_reserved_cus ||
      hipExtStreamCreateWithCUMask(&w.s2, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
    (void)hipGetLastError();
    w.s2 = nullptr;
    DUST_HIP(hipStreamCreateWithFlags(&w.s2, hipStreamNonBlocking));
  }
  DUST_HIP(hipEventCreateWithFlags(&w.ev_in, hipEventDisableTiming));
  DUST_HIP(hipEventCreateWithFlags(&w.ev_u12, hipEventDisableTiming));
  DUST_HIP(hipEventCreateWithFlags(&w.ev_rest, hipEventDisableTiming));
  DUST_HIP(hipEventCreateWithFlags(&w.ev_done, hipEventDisableTiming));
  DUST_TRY(w.hflag.reserve(4));
  return 0;
}

static void getrf_release() {
  GetrfWs &w = g_getrf;
  for (hipEvent_t *e : {&w.ev_in, &w.ev_u12, &w.ev_rest, &w.ev_done})
    if (*e) {
      (void)hipEventDestroy(*e);
      *e = nullptr;
    }
  for (hipStream_t *q : {&w.s1, &w.s2})
    if (*q) {
      (void)hipStreamDestroy(*q);
      *q = nullptr;
    }
  w.hflag.release();
}

int lu_getrf_dev(int m, int n, double *dA, long lda, int *dipiv, hipStream_t cst, int nb,
                 int *info_out) {
  DUST_CHECK(m >= 0 && n >= 0 && lda >= std::max(1, m), DUST_E_ARG, "invalid getrf dimensions");
  *info_out = 0;
  if (m == 0 || n == 0) return 0;
  nb = 64;  // the sub-panel width (the speculative block and k_panel's)
  GetrfWs &W = g_getrf;
  DUST_TRY(getrf_streams(W));
  hipStream_t s1 = W.s1, s2 = W.s2;
  const int cus = device_cus();
  const int Gmax = cus;
  const int mn = std::min(m, n);
  const int npanel = ceil_div(mn, nb);
  // workspace: barrier + candidates | speculation verdicts | original block |
  // original rows below it | L^-1 of every sub-panel's diagonal block
  size_t wbytes = 256 + sizeof(double) * 2 * Gmax + sizeof(int) * 2 * Gmax +
                  sizeof(double) * 2 * (size_t)Gmax * nb + sizeof(double) * 2 * nb + 64 +
                  sizeof(unsigned) * (Gmax + 16);
  wbytes = (wbytes + 255) & ~(size_t)255;
  const size_t off_fail = wbytes;
  wbytes += ((sizeof(int) * npanel + 255) & ~(size_t)255);
  const size_t off_d = wbytes;
  wbytes += sizeof(double) * SP * SP;
  const size_t off_w = wbytes;
  wbytes += sizeof(double) * (size_t)std::max(m, 1) * SP;
  const size_t off_li = wbytes;
  wbytes += sizeof(double) * SP * SP * (size_t)npanel;
  TmpDev<char> wsb;
  DUST_TRY(wsb.alloc(wbytes));
  char *ws = wsb.p;
  DUST_HIP(hipEventRecord(W.ev_in, cst));
  DUST_HIP(hipStreamWaitEvent(s1, W.ev_in, 0));
  DUST_HIP(hipMemsetAsync(ws, 0, off_d, s1));
  unsigned *bar = (unsigned *)ws;
  int *dinfo = (int *)(ws + 16);
  double *cval = (double *)(ws + 256);
  int *crow = (int *)(cval + 2 * Gmax);
  double *crowv = (double *)(crow + 2 * Gmax + 2);
  crowv = (double *)(((uintptr_t)crowv + 15) & ~(uintptr_t)15);
  double *topv = crowv + 2 * (size_t)Gmax * nb;
  unsigned *arr = (unsigned *)(topv + 2 * nb);
  int *spec_fail = (int *)(ws + off_fail);
  double *Dorig = (double *)(ws + off_d);
  double *Wl = (double *)(ws + off_w);
  double *Linv = (double *)(ws + off_li);  // [npanel][64 x 64]
  unsigned gen0 = 0;
  int rc = 0;
  bool s2_busy = false;  // ev_rest marks the last trailing update on s2
  auto launch_check = [&](const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("%s launch failed: %s", what, hipGetErrorString(e));
      rc = DUST_E_HIP;
    }
  };
  // factor the outer panel [K0, K0 + JB) on s1 (its columns up to date)
  auto factor_panel = [&](int K0, int JB) -> int {
    for (int k = K0; k < K0 + JB; k += nb) {
      const int jb = std::min(nb, K0 + JB - k);
      const int mrows = m - k;
      int *fail = spec_fail + (k / nb);
      // speculative sub-panel: a full 64-row block and at least one row below
      double *Li = Linv + (size_t)(k / nb) * SP * SP;
      bool exact = true;
      if (jb == nb && mrows > SP) {
        const int nrow = mrows - SP;
        k_spec_diag<<<1, 256, 0, s1>>>(dA, lda, k, Dorig, dipiv, Li, fail);
        k_spec_l21<<<ceil_div(nrow, 64), 64, 0, s1>>>(dA, lda, k, nrow, Wl, nrow, fail);
        launch_check("speculative panel");
        if (rc) return rc;
        DUST_HIP(hipMemcpyAsync(W.hflag.p, fail, sizeof(int), hipMemcpyDeviceToHost, s1));
        DUST_HIP(hipStreamSynchronize(s1));
        exact = W.hflag.p[0] != 0;
        if (exact) {
          const long tot = (long)SP * SP + (long)nrow * SP;
          k_spec_restore<<<ceil_div(tot, 256), 256, 0, s1>>>(dA, lda, k, nrow, Dorig, Wl, nrow);
        }
      }
      if (exact) {
        // all CUs for k_panel's resident grid: the trailing update first
        if (s2_busy) DUST_HIP(hipStreamWaitEvent(s1, W.ev_rest, 0));
        int G = std::min(std::min(Gmax, 128), std::max(1, (mrows + 63) / 64));
        int R = (mrows + G - 1) / G;
        G = (mrows + R - 1) / R;
        size_t lds = sizeof(double) * ((size_t)R * jb + 2 * jb);
        DUST_CHECK(lds <= 150 * 1024, DUST_E_UNSUPPORTED, "panel of %d x %d does not fit in LDS",
                   R, jb);
        PanelArgs P{dA, lda, m, k, jb, R, dipiv, dinfo, arr, bar + 2, gen0, cval, crow, crowv,
                    topv, nullptr};
        gen0 += jb;
        if (lds > 64 * 1024)
          (void)hipFuncSetAttribute((const void *)k_panel,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        // G <= 128 blocks of one CU each, all resident at once (the grid
        // barrier is bounded and reports a time-out otherwise).  An ordinary
        // launch: the kernel does its own synchronisation and needs none of
        // the cooperative-launch resources, which another process may hold.
        k_panel<<<G, 256, lds, s1>>>(P);
        k_trtri_unit<<<1, 256, 0, s1>>>(dA + k + lda * k, lda, jb, Li);
        launch_check("panel");
        if (rc) return rc;
      }
      // interchanges on the other columns of the outer panel
      if (JB > jb)
        k_laswp<<<ceil_div(JB, 256), 256, 0, s1>>>(dA + lda * K0, lda, JB, k - K0, k - K0 + jb, k,
                                                   jb, dipiv);
      const int wrest = K0 + JB - k - jb;
      if (wrest > 0) {
        k_trsm_inv<<<ceil_div(wrest, 64), 256, 0, s1>>>(Li, jb, dA + k + lda * (k + jb), lda,
                                                        wrest);
        gemm_sub(m - k - jb, wrest, jb, dA + (k + jb) + lda * k, lda, dA + k + lda * (k + jb), lda,
                 dA + (k + jb) + lda * (k + jb), lda, s1);
      }
      launch_check("sub-panel update");
      if (rc) return rc;
    }
    return 0;
  };
  rc = factor_panel(0, std::min(LU_NB, mn));
  for (int K0 = 0; K0 < mn && !rc; K0 += LU_NB) {
    const int JB = std::min(LU_NB, mn - K0);
    // the panel's interchanges on the columns outside it (after the last
    // trailing update)
    if (s2_busy) DUST_HIP(hipStreamWaitEvent(s1, W.ev_rest, 0));
    s2_busy = false;
    k_laswp<<<ceil_div(n, 256), 256, 0, s1>>>(dA, lda, n, K0, K0 + JB, K0, JB, dipiv);
    launch_check("laswp");
    if (rc) break;
    const int nrest = n - K0 - JB;
    if (nrest <= 0) break;
    // U12 = L11^-1 A12, blocked by the sub-panels: their inverse diagonal
    // blocks, then the update of the rows below by GEMM
    for (int b = 0; b < JB; b += nb) {
      const int jb = std::min(nb, JB - b);
      double *Bb = dA + (K0 + b) + lda * (K0 + JB);
      k_trsm_inv<<<ceil_div(nrest, 64), 256, 0, s1>>>(Linv + (size_t)((K0 + b) / nb) * SP * SP,
                                                      jb, Bb, lda, nrest);
      if (b + jb < JB)
        gemm_sub(JB - b - jb, nrest, jb, dA + (K0 + b + jb) + lda * (K0 + b), lda, Bb, lda,
                 Bb + jb, lda, s1);
    }
    launch_check("U12");
    if (rc) break;
    const int K1 = K0 + JB;
    const int JB2 = K1 < mn ? std::min(LU_NB, mn - K1) : 0;
    const int mrest = m - K1;
    // the next panel's columns first (s1), the rest on s2 meanwhile
    if (JB2 > 0)
      gemm_sub(mrest, JB2, JB, dA + K1 + lda * K0, lda, dA + K0 + lda * K1, lda,
               dA + K1 + lda * K1, lda, s1);
    if (nrest > JB2) {
      DUST_HIP(hipEventRecord(W.ev_u12, s1));
      DUST_HIP(hipStreamWaitEvent(s2, W.ev_u12, 0));
      gemm_sub(mrest, nrest - JB2, JB, dA + K1 + lda * K0, lda, dA + K0 + lda * (K1 + JB2), lda,
               dA + K1 + lda * (K1 + JB2), lda, s2);
      DUST_HIP(hipEventRecord(W.ev_rest, s2));
      s2_busy = true;
    }
    launch_check("trailing update");
    if (rc) break;
    if (JB2 > 0) rc = factor_panel(K1, JB2);
  }
  if (s2_busy) (void)hipStreamWaitEvent(s1, W.ev_rest, 0);
  unsigned hb[3] = {0, 0, 0};
  int hinfo = 0;
  if (!rc) {
    DUST_HIP(hipMemcpyAsync(hb, bar, sizeof(hb), hipMemcpyDeviceToHost, s1));
    DUST_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, s1));
  }
  DUST_HIP(hipStreamSynchronize(s1));
  DUST_HIP(hipStreamSynchronize(s2));
  // the caller's stream continues after the factorisation
  DUST_HIP(hipEventRecord(W.ev_done, s1));
  DUST_HIP(hipStreamWaitEvent(cst, W.ev_done, 0));
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
  getrf_release();
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
  gemm_sub(M, N, K, A, lda, B, ldb, C, ldc, st);
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
  // blocked: diagonal blocks of 64 (one thread per column), off-diagonal
  // updates by GEMM
  for (int k = 0; k < m; k += 64) {
    int jb = std::min(64, m - k);
    k_trsm_llnu<<<ceil_div(n, 64), 64, 0, st>>>(L + k + ldl * k, ldl, B + k, ldb, jb, n);
    if (k + jb < m) DUST_TRY(lu_gemm_sub_dev(m - k - jb, n, jb, L + (k + jb) + ldl * k, ldl, B + k, ldb,
                                             B + k + jb, ldb, st));
  }
  DUST_HIP(hipGetLastError());
  return 0;
}

}  // namespace dust

using namespace dust;

extern "C" {

int dust_lu_set_tuning(const char *name, int value) {
  DUST_CHECK(name, DUST_E_ARG, "null knob name");
  const std::string k(name);
  const int cus = device_cus();
  DUST_CHECK(value >= 0 && value <= cus / 2, DUST_E_ARG,
             "reserved compute units must be in 0..%d", cus / 2);
  if (k == "getrf_reserved_cus") {
    DUST_HIP(hipDeviceSynchronize());
    getrf_release();  // the lookahead streams are rebuilt with the new mask
    g_lu_reserved_cus = value;
    return 0;
  }
  if (k == "mixed_reserved_cus") return mixed_set_reserved_cus(value);
  set_error("unknown LU knob '%s'", name);
  return DUST_E_ARG;
}

int dust_lu_get_tuning(const char *name, int *value) {
  DUST_CHECK(name && value, DUST_E_ARG, "null argument");
  const std::string k(name);
  if (k == "getrf_reserved_cus") {
    *value = g_lu_reserved_cus;
    return 0;
  }
  if (k == "mixed_reserved_cus") {
    *value = mixed_reserved_cus();
    return 0;
  }
  set_error("unknown LU knob '%s'", name);
  return DUST_E_ARG;
}

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
