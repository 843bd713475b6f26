// Tuned mixed-precision dense solve A X = B for the AIC systems of
// src/linsys/mod_linsys.f90 (dgetrf 236/435 + dgetrs 473): the LU runs on the
// matrix cores in FP32 or in row-scaled FP16, the diagonal blocks are
// factored in FP64, and FP64 residuals drive iterative refinement.
//
// Factorisation (np = n rounded up to 128, padded with an identity block):
//   * W = R A in FP32, R a power-of-two row scaling (max |w_ij| per row in
//     [1/2, 1)) so that FP16 operands stay in range.
//   * Block columns of NB = 128.  The diagonal block is factored by one
//     workgroup in FP64 with row interchanges among its own 128 rows
//     (pivot = largest magnitude in the block column part).  The multipliers
//     of the rows below, L21 = A21 U11^-1, are then checked against a
//     threshold tau = TAU: max |l_ij| <= tau means the block pivot dominates the
//     whole column to within tau (threshold partial pivoting).  If any block
//     fails the check, or a pivot is zero, the solve is redone by the plain
//     FP64 partial-pivoting path.
//   * L21 and U12 come from MFMA triangular solves that use FP64-inverted
//     32x32 diagonal sub-blocks; the trailing update is delayed over DEPTH
//     block columns (K = DEPTH * 128 per pass over the trailing matrix) and
//     runs as an FP16 or FP32 MFMA GEMM with FP32 accumulation.  Operands are
//     written once per block column in MFMA fragment order.
//
// Refinement: x0 = U^-1 L^-1 P R b; r = b - A x in FP64; d = U^-1 L^-1 P R r;
// x += d, until ||r||_inf / (||A||_inf ||x||_inf) <= tol.  Triangular
// solves are one persistent launch each: block row kb waits on a per-block
// ready flag of every block row it depends on (generation-stamped, no reset).
//
// The workspace is process-wide and grows on demand; calls are not
// re-entrant.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "common.h"
#include "lu_internal.h"

namespace dust {
namespace {

constexpr int NB = 128;         // diagonal block
constexpr int SP = 32;          // sub-panel / inverted sub-block
constexpr int LDD = NB + 1;     // LDS leading dimension of the FP64 diagonal block
constexpr int DEPTH = 4;        // block columns per trailing update
constexpr int NR = 8;           // right-hand sides per solve pass
// Multiplier bound of the threshold pivoting: |l_ij| <= 10, the threshold
// u = 0.1 of sparse direct solvers (Duff, Erisman & Reid, "Direct Methods for
// Sparse Matrices", ch. 7).  Rotor AICs with thin blades reach |l| ~ 2.6
// below a diagonal block; refinement corrects the extra growth.
constexpr double TAU = 10.0;

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double rdlane(double v, int l) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  unsigned lo = __builtin_amdgcn_readlane((unsigned)u, l);
  unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Wavefront maximum of a 32-bit key by DPP row shifts and row broadcasts.
__device__ inline unsigned wave_umax(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}

// Pivot key: FP32 magnitude with its 7 low mantissa bits replaced by
// (127 - row), so the maximum picks the largest |a| and, among values equal
// to 2^-16 relative, the smallest row.
__device__ inline unsigned piv_key(double v, int row) {
  return (__float_as_uint(fabsf((float)v)) & ~127u) | (unsigned)(127 - row);
}

__device__ inline void wave_max_atomic(float v, unsigned *dst) {
  unsigned k = wave_umax(__float_as_uint(fabsf(v)));
  if ((threadIdx.x & 63) == 0) atomicMax(dst, k);
}

// ---------------------------------------------------------------------------
// Conversion: row maxima of |A|, then W = R A (FP32, padded identity).
// ---------------------------------------------------------------------------
__global__ void k_mx_fill_u32(unsigned *p, int n, unsigned v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Streaming over A (n x n, column-major) for the row maxima and the residual:
// a workgroup of 128 threads takes 256 rows (two per thread) and a chunk of
// `cw` columns, eight columns in flight per thread.  VEC = 1 (one 16-byte
// load per column and thread) only when A is 16-byte aligned and n and lda
// are even, so that both rows of every pair exist.
constexpr int SR = 256;  // rows per streaming workgroup
constexpr int SU = 8;    // columns in flight per thread

template <int VEC>
__device__ __forceinline__ double2 load_pair(const double *__restrict__ p, bool two) {
  if (VEC) return *(const double2 *)p;
  return make_double2(p[0], two ? p[1] : 0.0);
}

// Column chunks of the streaming kernels: about 2048 workgroups, chunks of
// 64..512 columns (multiple of 8).
inline int stream_chunk(int n, int *nch) {
  const int rb = ceil_div(n, SR);
  int c = std::max(1, std::min(ceil_div(n, 64), ceil_div(2048, rb)));
  int cw = std::min(512, std::max(64, ceil_div(ceil_div(n, c), SU) * SU));
  *nch = ceil_div(std::max(n, 1), cw);
  return cw;
}

inline bool stream_vec(const double *A, long lda, int n) {
  return ((uintptr_t)A & 15) == 0 && (lda & 1) == 0 && (n & 1) == 0;
}

// rowmax[i] = max over the chunk of |A(i, j)| (FP32 bits, atomic max).
template <int VEC>
__global__ void __launch_bounds__(128) k_mx_rowmax(const double *__restrict__ A, long lda, int n,
                                                   int cw, unsigned *rowmax) {
  const int i = blockIdx.x * SR + 2 * threadIdx.x;
  if (i >= n) return;
  const bool two = i + 1 < n;
  const int j0 = blockIdx.y * cw, nj = min(n - j0, cw);
  const double *Ap = A + i + lda * (long)j0;
  double m0 = 0.0, m1 = 0.0;
  int j = 0;
  for (; j + SU <= nj; j += SU) {
    double2 a[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) a[u] = load_pair<VEC>(Ap + lda * (long)(j + u), two);
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      m0 = fmax(m0, fabs(a[u].x));
      m1 = fmax(m1, fabs(a[u].y));
    }
  }
  for (; j < nj; ++j) {
    double2 a = load_pair<0>(Ap + lda * (long)j, two);
    m0 = fmax(m0, fabs(a.x));
    m1 = fmax(m1, fabs(a.y));
  }
  atomicMax(&rowmax[i], __float_as_uint((float)m0));
  if (two) atomicMax(&rowmax[i + 1], __float_as_uint((float)m1));
}

__global__ void k_mx_rowscale(const unsigned *rowmax, int n, int np, int scaled, double *rs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double s = 1.0;
  if (scaled && i < n) {
    float m = __uint_as_float(rowmax[i]);
    if (m > 0.f && isfinite(m)) {
      int e;
      (void)frexpf(m, &e);  // m = f 2^e, f in [1/2, 1)
      s = ldexp(1.0, -e);
    }
  }
  rs[i] = s;
}

__global__ void k_mx_convert(const double *__restrict__ A, long lda, int n, int np,
                             const double *__restrict__ rs, float *W) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j0 = blockIdx.y * 4;
  if (i >= np) return;
  double s = rs[i];
  for (int j = j0; j < j0 + 4; ++j) {
    float v;
    if (i < n && j < n) v = (float)(A[i + lda * j] * s);
    else v = (i == j) ? 1.f : 0.f;
    W[i + (long)np * j] = v;
  }
}

// ---------------------------------------------------------------------------
// Diagonal block: FP64 LU with interchanges among the block's rows, plus the
// FP64 inverses of the four 32x32 diagonal sub-blocks of L and U.
// ---------------------------------------------------------------------------
struct DiagArgs {
  float *W;
  long ld;
  int k;          // first row/column of the block
  int *piv;       // [np]: global 0-based pivot row of each row (LAPACK order)
  int *perm;      // [np]: factored row -> original row
  int *npiv;      // [nblk]: interchanges in the block
  float *dinv;    // [nblk][8][32*32]: Linv sub-blocks 0..3, Uinv sub-blocks 4..7
  unsigned *flags;  // [0] max |l| bits, [1] zero pivot
};

// Wave 0: factor columns [c0, c0+32) of the block, rows [c0, 128), held in
// registers: lane l holds the slots of rows l and l + 64 at the start.  Rows
// are not moved while the sub-panel is factored: each slot carries the row
// position of its data (pos0, pos1), an interchange of positions g and p
// only exchanges two labels, and every slot is stored at its final position
// at the end.  The pivot of column g is the largest |a| among positions >= g
// (smallest position on ties); the interchanges are recorded LAPACK-style
// in spiv.  One template instance per column keeps every register index a
// compile-time constant.
template <int J>
__device__ __forceinline__ void subpanel_step(double (&a0)[SP], double (&a1)[SP], int &pos0,
                                              int &pos1, int c0, int lane, int *spiv,
                                              unsigned *flags) {
  const int g = c0 + J;
  const unsigned k0 = pos0 >= g ? piv_key(a0[J], pos0) : 0u;
  const unsigned k1 = pos1 >= g ? piv_key(a1[J], pos1) : 0u;
  const int p = 127 - (int)(wave_umax(max(k0, k1)) & 127u);
  // the slot holding position p (wave-uniform)
  const unsigned long long m0 = __ballot(pos0 == p);
  const unsigned long long m1 = __ballot(pos1 == p);
  const int ls = __ffsll((long long)(m0 ? m0 : m1)) - 1;
  double u[SP];
  if (m0) {
#pragma unroll
    for (int c = J; c < SP; ++c) u[c] = rdlane(a0[c], ls);
  } else {
#pragma unroll
    for (int c = J; c < SP; ++c) u[c] = rdlane(a1[c], ls);
  }
  const double piv = u[J];
  if (lane == 0) {
    spiv[g] = p;
    if (piv == 0.0) atomicOr(&flags[1], 1u);
  }
  pos0 = pos0 == g ? p : (pos0 == p ? g : pos0);
  pos1 = pos1 == g ? p : (pos1 == p ? g : pos1);
  // 1 / piv: hardware reciprocal and two Newton steps
  double rp = __builtin_amdgcn_rcp(piv);
  rp = fma(fma(-piv, rp, 1.0), rp, rp);
  rp = fma(fma(-piv, rp, 1.0), rp, rp);
  // multipliers and column J + 1 (the next pivot column) first
  const double l0 = pos0 > g ? a0[J] * rp : 0.0;
  const double l1 = pos1 > g ? a1[J] * rp : 0.0;
  if (pos0 > g) a0[J] = l0;
  if (pos1 > g) a1[J] = l1;
  if constexpr (J + 1 < SP) {
    a0[J + 1] = fma(-l0, u[J + 1], a0[J + 1]);
    a1[J + 1] = fma(-l1, u[J + 1], a1[J + 1]);
  }
#pragma unroll
  for (int c = J + 2; c < SP; ++c) {
    a0[c] = fma(-l0, u[c], a0[c]);
    a1[c] = fma(-l1, u[c], a1[c]);
  }
}

template <int... Js>
__device__ __forceinline__ void subpanel_steps(std::integer_sequence<int, Js...>, double (&a0)[SP],
                                               double (&a1)[SP], int &pos0, int &pos1, int c0,
                                               int lane, int *spiv, unsigned *flags) {
  (subpanel_step<Js>(a0, a1, pos0, pos1, c0, lane, spiv, flags), ...);
}

// sorig[q]: original row (within the block) of the row now at position q.
__device__ __forceinline__ void diag_subpanel(double *sD, int *sorig, int c0, int lane, int *spiv,
                                              unsigned *flags) {
  double a0[SP], a1[SP];
#pragma unroll
  for (int j = 0; j < SP; ++j) {
    a0[j] = sD[(c0 + j) * LDD + lane];
    a1[j] = sD[(c0 + j) * LDD + lane + 64];
  }
  const int o0 = sorig[lane], o1 = sorig[lane + 64];
  int pos0 = lane, pos1 = lane + 64;
  subpanel_steps(std::make_integer_sequence<int, SP>{}, a0, a1, pos0, pos1, c0, lane, spiv, flags);
#pragma unroll
  for (int j = 0; j < SP; ++j) {
    sD[(c0 + j) * LDD + pos0] = a0[j];
    sD[(c0 + j) * LDD + pos1] = a1[j];
  }
  sorig[pos0] = o0;
  sorig[pos1] = o1;
}

// BATCH: workgroup b factors the block of items[b] (one block per system).
template <int BATCH>
__global__ void __launch_bounds__(256) k_mx_diag(DiagArgs one, const DiagArgs *__restrict__ items) {
  const DiagArgs a = BATCH ? items[blockIdx.x] : one;
  extern __shared__ double sD[];  // NB x LDD, column-major
  __shared__ int spiv[NB];
  __shared__ int sorig[NB];
  __shared__ double sLi[SP * SP];  // inverse of the sub-panel's unit-lower L11, column-major
  __shared__ double srd[NB];       // 1 / U(i, i)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // the block, 16 bytes per load, all loads in flight before the LDS stores
  // (W and its leading dimension are multiples of 128 floats)
  {
    const float *Wd = a.W + a.k + a.ld * a.k;
    float4 v[NB * NB / 4 / 256];
#pragma unroll
    for (int it = 0; it < NB * NB / 4 / 256; ++it) {
      const int q = tid + 256 * it, c = q >> 5, r = (q & 31) * 4;
      v[it] = *(const float4 *)(Wd + r + a.ld * c);
    }
#pragma unroll
    for (int it = 0; it < NB * NB / 4 / 256; ++it) {
      const int q = tid + 256 * it, c = q >> 5, r = (q & 31) * 4;
      double *d = sD + c * LDD + r;
      d[0] = v[it].x;
      d[1] = v[it].y;
      d[2] = v[it].z;
      d[3] = v[it].w;
    }
  }
  if (tid < NB) sorig[tid] = tid;
  __syncthreads();
  for (int sp = 0; sp < NB / SP; ++sp) {
    const int c0 = sp * SP;
    if (wv == 0) diag_subpanel(sD, sorig, c0, lane, spiv, a.flags);
    __syncthreads();
    // interchanges of this sub-panel on the other 96 columns; meanwhile wave 3
    // inverts the sub-panel's L11 (final now: later interchanges only touch
    // rows below it) blockwise: L11 = [A 0; C D] (16 x 16 blocks), inverse
    // [A^-1 0; -D^-1 C A^-1  D^-1]; lanes 0-15 / 16-31 take a column of A^-1 /
    // D^-1 by forward substitution, then two FP64 MFMA products
    if (wv == 3) {
      const int h = lane >> 4, c = lane & 15;
      if (lane < 32) {
        const int b0 = c0 + 16 * h;
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = (i == c) ? 1.0 : 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const double *col = sD + (b0 + i) * LDD + b0;  // unit lower, rows i+1..
#pragma unroll
          for (int r = i + 1; r < 16; ++r) x[r] = fma(-col[r], x[i], x[r]);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) sLi[(16 * h + c) * SP + 16 * h + i] = x[i];
      }
      // T = C A^-1, then the lower-left block -D^-1 T (f64 16x16x4 layouts:
      // A operand row lane % 16, k lane / 16; B operand k lane / 16, column
      // lane % 16; output column lane % 16, rows 4 i + lane / 16)
      d4 t = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 16; kk += 4) {
        const double av = sD[(c0 + kk + (lane >> 4)) * LDD + c0 + 16 + (lane & 15)];  // C
        const double bv = sLi[(lane & 15) * SP + kk + (lane >> 4)];                   // A^-1
        t = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, t, 0, 0, 0);
      }
      // T to the (still zero) upper-right block of sLi as scratch, then -D^-1 T
#pragma unroll
      for (int i = 0; i < 4; ++i) sLi[(16 + (lane & 15)) * SP + 4 * i + (lane >> 4)] = t[i];
      d4 xo = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 16; kk += 4) {
        const double av = -sLi[(16 + kk + (lane >> 4)) * SP + 16 + (lane & 15)];  // -D^-1
        const double bv = sLi[(16 + (lane & 15)) * SP + kk + (lane >> 4)];        // T (k, col)
        xo = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, xo, 0, 0, 0);
      }
      // lower-left block (rows 16.., columns ..15); the upper-right block is zero
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        sLi[(lane & 15) * SP + 16 + 4 * i + (lane >> 4)] = xo[i];
        sLi[(16 + (lane & 15)) * SP + 4 * i + (lane >> 4)] = 0.0;
      }
      float *Li = a.dinv + ((long)(a.k / NB) * 8 + sp) * SP * SP;
      for (int q = lane; q < SP * SP; q += 64) Li[q] = (float)sLi[q];
    }
    if (tid < NB - SP) {
      const int col = tid < c0 ? tid : tid + SP;
      double *cp = sD + col * LDD;
      for (int i = 0; i < SP; ++i) {
        int p = spiv[c0 + i];
        if (p != c0 + i) {
          double t = cp[c0 + i];
          cp[c0 + i] = cp[p];
          cp[p] = t;
        }
      }
    }
    __syncthreads();
    const int m2 = NB - c0 - SP;  // rows/columns after the sub-panel
    if (m2 == 0) break;
    // U12 of the sub-panel = L11^-1 A12 on FP64 MFMA; a wave owns whole
    // 16-column blocks (it reads all 32 rows of a block before writing it)
    for (int tj = wv; tj < m2 / 16; tj += 4) {
      const int cc0 = c0 + SP + 16 * tj;
      d4 acc0 = (d4){0.0, 0.0, 0.0, 0.0}, acc1 = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < SP; kk += 4) {
        const double bv = sD[(cc0 + (lane & 15)) * LDD + c0 + kk + (lane >> 4)];
        const double av0 = sLi[(kk + (lane >> 4)) * SP + (lane & 15)];
        const double av1 = sLi[(kk + (lane >> 4)) * SP + 16 + (lane & 15)];
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv, acc1, 0, 0, 0);
      }
      // f64 16x16x4 output: column lane % 16, rows 4 i + lane / 16
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        sD[(cc0 + (lane & 15)) * LDD + c0 + 4 * i + (lane >> 4)] = acc0[i];
        sD[(cc0 + (lane & 15)) * LDD + c0 + 16 + 4 * i + (lane >> 4)] = acc1[i];
      }
    }
    __syncthreads();
    // trailing block: D22 -= L21 U12 on FP64 MFMA, 16x16 tiles over the waves
    {
      const int nt = m2 / 16;
      const int base = c0 + SP;
      for (int t = wv; t < nt * nt; t += 4) {
        const int ti = t % nt, tj = t / nt;
        const int r0 = base + 16 * ti, cc0 = base + 16 * tj;
        d4 acc;
        // f64 16x16x4 output: column lane % 16, rows 4 i + lane / 16
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = sD[(cc0 + (lane & 15)) * LDD + r0 + 4 * i + (lane >> 4)];
#pragma unroll
        for (int kk = 0; kk < SP; kk += 4) {
          double av = -sD[(c0 + kk + (lane >> 4)) * LDD + r0 + (lane & 15)];
          double bv = sD[(cc0 + (lane & 15)) * LDD + c0 + kk + (lane >> 4)];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) sD[(cc0 + (lane & 15)) * LDD + r0 + 4 * i + (lane >> 4)] = acc[i];
      }
    }
    __syncthreads();
  }
  // inverses of the four 32x32 diagonal sub-blocks of U (those of L were
  // formed after each sub-panel): waves 0 and 1, two sub-blocks each, column
  // c by back substitution in column order with the diagonal reciprocals
  if (tid < NB) srd[tid] = 1.0 / sD[tid * LDD + tid];
  __syncthreads();
  if (wv < 2) {
    const int sb = 2 * wv + (lane >> 5), c = lane & 31, b0 = sb * SP;
    float *Ui = a.dinv + ((long)(a.k / NB) * 8 + 4 + sb) * SP * SP;
    double x[SP];
#pragma unroll
    for (int i = 0; i < SP; ++i) x[i] = (i == c) ? 1.0 : 0.0;
#pragma unroll
    for (int i = SP - 1; i >= 0; --i) {
      const double *col = sD + (b0 + i) * LDD + b0;  // U(0..i, i)
      x[i] *= srd[b0 + i];
#pragma unroll
      for (int r = 0; r < i; ++r) x[r] = fma(-col[r], x[i], x[r]);
    }
#pragma unroll
    for (int i = 0; i < SP; ++i) Ui[i + SP * c] = (float)x[i];
  }
  // pivots, permutation (sorig), interchange count, write-back
  if (tid < NB) {
    a.piv[a.k + tid] = a.k + spiv[tid];
    a.perm[a.k + tid] = a.k + sorig[tid];
  }
  if (wv == 0) {
    const unsigned long long b0 = __ballot(spiv[lane] != lane);
    const unsigned long long b1 = __ballot(spiv[lane + 64] != lane + 64);
    if (lane == 0) a.npiv[a.k / NB] = __popcll(b0) + __popcll(b1);
  }
  {
    float *Wo = a.W + a.k + a.ld * a.k;
#pragma unroll 4
    for (int it = 0; it < NB * NB / 4 / 256; ++it) {
      const int q = tid + 256 * it, c = q >> 5, r = (q & 31) * 4;
      const double *d = sD + c * LDD + r;
      *(float4 *)(Wo + r + a.ld * c) = make_float4((float)d[0], (float)d[1], (float)d[2], (float)d[3]);
    }
  }
}

// Interchanges of the diagonal block at row k on every column outside the
// block.  BATCH: items[i] owns workgroups [starts[i], starts[i+1]).
struct LaswpItem {
  float *W;
  long ld;
  int np, k;
  const int *piv, *npiv;
};

template <int BATCH>
__global__ void k_mx_laswp(LaswpItem one, const LaswpItem *__restrict__ items,
                           const int *__restrict__ starts, int m) {
  int bx = blockIdx.x;
  LaswpItem a = one;
  if (BATCH) {
    const int i = batch_item(starts, m, bx);
    a = items[i];
    bx -= starts[i];
  }
  const int k = a.k;
  if (a.npiv[k / NB] == 0) return;
  int c = bx * blockDim.x + threadIdx.x;
  if (c >= a.np || (c >= k && c < k + NB)) return;
  float *col = a.W + a.ld * c;
  const int *piv = a.piv;
  for (int i = k; i < k + NB; ++i) {
    int p = piv[i];
    if (p != i) {
      float t = col[i];
      col[i] = col[p];
      col[p] = t;
    }
  }
}

// ---------------------------------------------------------------------------
// Block-column / block-row triangular solves on FP32 MFMA.
//   L tiles: X U11 = A21 for 64 rows (X -> W and L fragments), max |x|.
//   U tiles: L11 Y = A12 for 64 columns (Y -> W and U fragments).
// mode 1 (inverse): A21 = I / A12 = I, output to the 128x128 Uinv / Linv of
// the block instead (used by the refinement solves).
// Fragment order: FP16 (E = 8, KS = 32): frag (tile, s) lane l holds
// rows/cols 16 tile + l % 16 and k = 32 s + 8 (l / 16) + e; FP32 (E = 1,
// KS = 4): k = 4 s + l / 16.
// ---------------------------------------------------------------------------
constexpr int LDX = 64 + 1;   // LDS leading dimension of the 64-wide tile (floats)
constexpr int LDT = NB + 1;   // LDS leading dimension of the 128x128 triangle

// Column-major FP32 tile (ROWS x COLS, ROWS a multiple of 4; source and
// leading dimension 16-byte aligned) -> LDS with leading dimension sld, by a
// 256-thread workgroup: 16-byte loads, eight in flight per thread before the
// LDS stores.
template <int ROWS, int COLS>
__device__ __forceinline__ void tile_to_lds(const float *__restrict__ src, long ld, float *dst,
                                            int sld, int tid) {
  constexpr int R4 = ROWS / 4, PER = ROWS * COLS / 4 / 256, B = PER < 8 ? PER : 8;
  static_assert(ROWS % 4 == 0 && (ROWS * COLS / 4) % 256 == 0 && PER % B == 0, "tile shape");
#pragma unroll
  for (int b0 = 0; b0 < PER; b0 += B) {
    float4 v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int q = tid + 256 * (b0 + u), c = q / R4, r = (q % R4) * 4;
      v[u] = *(const float4 *)(src + r + ld * c);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int q = tid + 256 * (b0 + u), c = q / R4, r = (q % R4) * 4;
      float *d = dst + c * sld + r;
      d[0] = v[u].x;
      d[1] = v[u].y;
      d[2] = v[u].z;
      d[3] = v[u].w;
    }
  }
}

// 4 x 32 x 32 inverse sub-blocks (contiguous, 16-byte aligned) -> LDS.
__device__ __forceinline__ void inv_to_lds(const float *__restrict__ src, float *dst, int tid) {
  float4 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) v[u] = ((const float4 *)src)[tid + 256 * u];
#pragma unroll
  for (int u = 0; u < 4; ++u) ((float4 *)dst)[tid + 256 * u] = v[u];
}

struct TrsmArgs {
  float *W;
  long ld;
  int k, np, nl;    // nl: number of L tiles (the rest are U tiles)
  int f16;          // fragment precision
  int mode;         // 0: factorisation, 1: block inverses
  const float *dinv;
  void *Lf, *Uf;    // fragment panels of this block column
  unsigned *flags;
  float *Linv, *Uinv;  // [nblk][128*128] (mode 1)
};

__device__ inline void frag_store_L(void *Lf, int f16, int rt, int lane, const float *sX, int rloc) {
  // rows rloc.. of sX (column-major, LDX) -> fragments of row tile rt
  if (f16) {
    h8 *F = (h8 *)Lf;
    for (int s = 0; s < NB / 32; ++s) {
      h8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        v[e] = (_Float16)sX[(32 * s + 8 * (lane >> 4) + e) * LDX + rloc + (lane & 15)];
      F[((long)rt * (NB / 32) + s) * 64 + lane] = v;
    }
  } else {
    float *F = (float *)Lf;
    for (int s = 0; s < NB / 4; ++s)
      F[((long)rt * (NB / 4) + s) * 64 + lane] = sX[(4 * s + (lane >> 4)) * LDX + rloc + (lane & 15)];
  }
}

__device__ inline void frag_store_U(void *Uf, int f16, int ct, int lane, const float *sY, int cloc) {
  // sY: rows 0..127 of columns cloc.. (column-major, LDT)
  if (f16) {
    h8 *F = (h8 *)Uf;
    for (int s = 0; s < NB / 32; ++s) {
      h8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        v[e] = (_Float16)sY[(cloc + (lane & 15)) * LDT + 32 * s + 8 * (lane >> 4) + e];
      F[((long)ct * (NB / 32) + s) * 64 + lane] = v;
    }
  } else {
    float *F = (float *)Uf;
    for (int s = 0; s < NB / 4; ++s)
      F[((long)ct * (NB / 4) + s) * 64 + lane] = sY[(cloc + (lane & 15)) * LDT + 4 * s + (lane >> 4)];
  }
}

// BATCH: items[i] owns workgroups [starts[i], starts[i+1]), numbered as the
// single launch's grid: 2 nl tiles (mode 0) or 4 x nblk (mode 1, x fastest).
template <int BATCH>
__global__ void __launch_bounds__(256) k_mx_trsm(TrsmArgs one, const TrsmArgs *__restrict__ items,
                                                 const int *__restrict__ starts, int m) {
  extern __shared__ float sm[];
  int bx = blockIdx.x, by = blockIdx.y;
  TrsmArgs a = one;
  if (BATCH) {
    const int i = batch_item(starts, m, bx);
    a = items[i];
    const int loc = bx - starts[i];
    bx = a.mode ? (loc & 3) : loc;
    by = a.mode ? (loc >> 2) : 0;
  }
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int k = a.mode ? by * NB : a.k;
  const int kb = k / NB;
  const float *Dg = a.W + k + a.ld * k;   // diagonal block
  const float *inv = a.dinv + (long)kb * 8 * SP * SP;
  if (bx < a.nl) {
    // ---------------- L tile: rows R0 .. R0+64 ----------------
    float *sX = sm;                   // 128 cols x LDX
    float *sU = sX + NB * LDX;        // 128 x LDT (upper triangle used)
    float *sI = sU + NB * LDT;        // 4 x 32 x 32 (Uinv sub-blocks)
    const int R0 = (a.mode == 0) ? k + NB + 64 * bx : 64 * bx;
    if (a.mode == 0) {
      tile_to_lds<64, NB>(a.W + R0 + a.ld * k, a.ld, sX, LDX, tid);
    } else {
      for (int q = tid; q < NB * 64; q += 256) {
        int r = q & 63, c = q >> 6;
        sX[c * LDX + r] = (R0 + r == c) ? 1.f : 0.f;
      }
    }
    tile_to_lds<NB, NB>(Dg, a.ld, sU, LDT, tid);
    inv_to_lds(inv + 4 * SP * SP, sI, tid);
    __syncthreads();
    const int r16 = 16 * wv;  // the wave's 16 rows
    for (int s = 0; s < NB / SP; ++s) {
      f4 acc[2];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          acc[t][i] = sX[(SP * s + 16 * t + (lane & 15)) * LDX + r16 + 4 * (lane >> 4) + i];
      for (int kk = 0; kk < SP * s; kk += 4) {
        float av = -sX[(kk + (lane >> 4)) * LDX + r16 + (lane & 15)];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          float bv = sU[(SP * s + 16 * t + (lane & 15)) * LDT + kk + (lane >> 4)];
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[t], 0, 0, 0);
        }
      }
      wsync();
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          sX[(SP * s + 16 * t + (lane & 15)) * LDX + r16 + 4 * (lane >> 4) + i] = acc[t][i];
      wsync();
      f4 out[2] = {(f4){0.f, 0.f, 0.f, 0.f}, (f4){0.f, 0.f, 0.f, 0.f}};
      const float *Ui = sI + s * SP * SP;
#pragma unroll
      for (int kk = 0; kk < SP; kk += 4) {
        float av = sX[(SP * s + kk + (lane >> 4)) * LDX + r16 + (lane & 15)];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          float bv = Ui[(kk + (lane >> 4)) + SP * (16 * t + (lane & 15))];
          out[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, out[t], 0, 0, 0);
        }
      }
      wsync();
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          sX[(SP * s + 16 * t + (lane & 15)) * LDX + r16 + 4 * (lane >> 4) + i] = out[t][i];
      wsync();
    }
    // outputs of the wave's 16 rows
    if (a.mode == 0) {
      float m = 0.f;
      bool nan = false;
      for (int q = lane; q < 16 * NB; q += 64) {
        int r = q & 15, c = q >> 4;
        float v = sX[c * LDX + r16 + r];
        m = fmaxf(m, fabsf(v));
        nan |= !(v == v);
        a.W[R0 + r16 + r + a.ld * (k + c)] = v;
      }
      wave_max_atomic(nan ? INFINITY : m, &a.flags[0]);
      frag_store_L(a.Lf, a.f16, (R0 + r16) / 16, lane, sX, r16);
    } else {
      float *Ui = a.Uinv + (long)kb * NB * NB;  // row-major
      for (int q = lane; q < 16 * NB; q += 64) {
        int c = q & (NB - 1), r = q >> 7;
        Ui[(R0 + r16 + r) * NB + c] = sX[c * LDX + r16 + r];
      }
    }
  } else {
    // ---------------- U tile: columns C0 .. C0+64 ----------------
    float *sY = sm;                   // 64 cols x LDT (rows 0..127)
    float *sL = sY + 64 * LDT;        // 128 x LDT (strict lower triangle used)
    float *sI = sL + NB * LDT;        // 4 x 32 x 32 (Linv sub-blocks)
    const int ub = bx - a.nl;
    const int C0 = (a.mode == 0) ? k + NB + 64 * ub : 64 * ub;
    if (a.mode == 0) {
      tile_to_lds<NB, 64>(a.W + k + a.ld * C0, a.ld, sY, LDT, tid);
    } else {
      for (int q = tid; q < NB * 64; q += 256) {
        int r = q & (NB - 1), c = q >> 7;
        sY[c * LDT + r] = (r == C0 + c) ? 1.f : 0.f;
      }
    }
    tile_to_lds<NB, NB>(Dg, a.ld, sL, LDT, tid);
    inv_to_lds(inv, sI, tid);
    __syncthreads();
    const int c16 = 16 * wv;  // the wave's 16 columns
    for (int s = 0; s < NB / SP; ++s) {
      f4 acc[2];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          acc[t][i] = sY[(c16 + (lane & 15)) * LDT + SP * s + 16 * t + 4 * (lane >> 4) + i];
      for (int kk = 0; kk < SP * s; kk += 4) {
        float bv = sY[(c16 + (lane & 15)) * LDT + kk + (lane >> 4)];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          float av = -sL[(kk + (lane >> 4)) * LDT + SP * s + 16 * t + (lane & 15)];
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[t], 0, 0, 0);
        }
      }
      wsync();
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          sY[(c16 + (lane & 15)) * LDT + SP * s + 16 * t + 4 * (lane >> 4) + i] = acc[t][i];
      wsync();
      f4 out[2] = {(f4){0.f, 0.f, 0.f, 0.f}, (f4){0.f, 0.f, 0.f, 0.f}};
      const float *Li = sI + s * SP * SP;
#pragma unroll
      for (int kk = 0; kk < SP; kk += 4) {
        float bv = sY[(c16 + (lane & 15)) * LDT + SP * s + kk + (lane >> 4)];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          float av = Li[(16 * t + (lane & 15)) + SP * (kk + (lane >> 4))];
          out[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, out[t], 0, 0, 0);
        }
      }
      wsync();
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          sY[(c16 + (lane & 15)) * LDT + SP * s + 16 * t + 4 * (lane >> 4) + i] = out[t][i];
      wsync();
    }
    if (a.mode == 0) {
      for (int q = lane; q < 16 * NB; q += 64) {
        int r = q & (NB - 1), c = q >> 7;
        a.W[k + r + a.ld * (C0 + c16 + c)] = sY[(c16 + c) * LDT + r];
      }
      frag_store_U(a.Uf, a.f16, (C0 + c16) / 16, lane, sY, c16);
    } else {
      float *Li = a.Linv + (long)kb * NB * NB;  // row-major
      for (int q = lane; q < 16 * NB; q += 64) {
        int r = q & (NB - 1), c = q >> 7;
        Li[r * NB + C0 + c16 + c] = sY[(c16 + c) * LDT + r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Update GEMM: W[r0.., c0..] -= sum_q Lfrag_q Ufrag_q over P block columns.
// Block tile 128 x 128, four waves of 64 x 64 (4 x 4 MFMA tiles).
// ---------------------------------------------------------------------------
struct GemmItem {
  float *W;
  long ld;
  int r0, c0;   // first row / column of the updated region
  int nrt, P;   // 128-row tiles of the region; block columns summed
  int ntiles;   // tiles of the region (the persistent single launch)
  const void *Lf, *Uf;
  long pbytes;  // bytes per fragment panel
};

// One 128 x 128 tile (bx, by) of the region.
template <int F16>
__device__ __forceinline__ void gemm_tile(const GemmItem &g, int bx, int by) {
  float *W = g.W;
  const long ld = g.ld, pbytes = g.pbytes;
  const void *Lf = g.Lf, *Uf = g.Uf;
  const int P = g.P;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int rt0 = (g.r0 + 128 * bx + 64 * (wv & 1)) / 16;
  const int ct0 = (g.c0 + 128 * by + 64 * (wv >> 1)) / 16;
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
  if (F16) {
    // fragments of step t = 4 q + s (block column q, 32-wide k slice s) read
    // three steps ahead into a ring of four register sets; the last steps
    // re-read step T - 1 (clamped index, no branches between the loads)
    const int T = P * (NB / 32);
    auto load = [&](int t, h8 (&av)[4], h8 (&bv)[4]) {
      t = min(t, T - 1);
      const int q = t >> 2, s = t & 3;
      const h8 *La = (const h8 *)((const char *)Lf + q * pbytes);
      const h8 *Ub = (const h8 *)((const char *)Uf + q * pbytes);
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = La[((long)(rt0 + i) * (NB / 32) + s) * 64 + lane];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Ub[((long)(ct0 + j) * (NB / 32) + s) * 64 + lane];
    };
    auto mma = [&](const h8 (&av)[4], const h8 (&bv)[4]) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[i], bv[j], acc[i][j], 0, 0, 0);
    };
    h8 a0[4], b0[4], a1[4], b1[4], a2[4], b2[4], a3[4], b3[4];
    load(0, a0, b0);
    load(1, a1, b1);
    load(2, a2, b2);
    for (int t = 0; t < T; t += 4) {
      load(t + 3, a3, b3);
      mma(a0, b0);
      load(t + 4, a0, b0);
      mma(a1, b1);
      load(t + 5, a1, b1);
      mma(a2, b2);
      load(t + 6, a2, b2);
      mma(a3, b3);
    }
  } else {
    for (int q = 0; q < P; ++q) {
      const float *La = (const float *)((const char *)Lf + q * pbytes);
      const float *Ub = (const float *)((const char *)Uf + q * pbytes);
#pragma unroll 4
      for (int s = 0; s < NB / 4; ++s) {
        float av[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) av[i] = La[((long)(rt0 + i) * (NB / 4) + s) * 64 + lane];
#pragma unroll
        for (int j = 0; j < 4; ++j) bv[j] = Ub[((long)(ct0 + j) * (NB / 4) + s) * 64 + lane];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
      }
    }
  }
  // output layout: row 4 (lane / 16) + i, column lane % 16; all sixteen
  // loads of the tile in flight before its stores
  f4 cv[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long row = 16L * (rt0 + i) + 4 * (lane >> 4);
      const long col = 16L * (ct0 + j) + (lane & 15);
      cv[i][j] = *(const f4 *)(W + row + ld * col);
    }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long row = 16L * (rt0 + i) + 4 * (lane >> 4);
      const long col = 16L * (ct0 + j) + (lane & 15);
      *(f4 *)(W + row + ld * col) = cv[i][j] - acc[i][j];
    }
}

// BATCH: items[i] owns workgroups [starts[i], starts[i+1]), tile (x, y) =
// (local % nrt, local / nrt).  Otherwise the workgroups take tiles t =
// blockIdx.x, blockIdx.x + gridDim.x, ... of the ntiles, (x, y) = (t % nrt,
// t / nrt), so that a launch can be held to a given number of workgroups.
template <int F16, int BATCH>
__global__ void __launch_bounds__(256) k_mx_gemm(GemmItem one, const GemmItem *__restrict__ items,
                                                 const int *__restrict__ starts, int m) {
  if (BATCH) {
    const int i = batch_item(starts, m, blockIdx.x);
    const GemmItem g = items[i];
    const int loc = blockIdx.x - starts[i];
    gemm_tile<F16>(g, loc % g.nrt, loc / g.nrt);
  } else {
    for (int t = blockIdx.x; t < one.ntiles; t += gridDim.x) gemm_tile<F16>(one, t % one.nrt, t / one.nrt);
  }
}

// Two regions in one launch (a block column and a block row of the same
// sub-panels): workgroups [0, a.ntiles) take the tiles of a, the rest those
// of b.
template <int F16>
__global__ void __launch_bounds__(256) k_mx_gemm2(GemmItem a, GemmItem b) {
  const int t = blockIdx.x;
  if (t < a.ntiles) gemm_tile<F16>(a, t % a.nrt, t / a.nrt);
  else if (t - a.ntiles < b.ntiles) gemm_tile<F16>(b, (t - a.ntiles) % b.nrt, (t - a.ntiles) / b.nrt);
}

// ---------------------------------------------------------------------------
// Refinement pieces.
// ---------------------------------------------------------------------------
// y[i][q] = (float)(rs[perm i] * r[perm i, q]); rows >= n and q >= nrhs -> 0.
__global__ void k_mx_gather(const double *__restrict__ r, long ldr, int n, int np, int nrhs,
                            const int *__restrict__ perm, const double *__restrict__ rs, float *y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float v[NR];
  int p = i < n ? perm[i] : 0;
  double s = i < n ? rs[p] : 0.0;
#pragma unroll
  for (int q = 0; q < NR; ++q) v[q] = (i < n && q < nrhs) ? (float)(s * r[p + ldr * q]) : 0.f;
  f4 *o = (f4 *)(y + (long)i * NR);
  o[0] = (f4){v[0], v[1], v[2], v[3]};
  o[1] = (f4){v[4], v[5], v[6], v[7]};
}

// x[i, q] (+)= y[i][q]
__global__ void k_mx_update_x(double *x, long ldx, int n, const float *__restrict__ y,
                              int accumulate) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < NR; ++q) {
    double d = (double)y[(long)i * NR + q];
    x[i + ldx * q] = accumulate ? x[i + ldx * q] + d : d;
  }
}

// Partial products of r = b - A x over column chunks (deterministic: the
// chunks are summed in order by k_mx_resid_fin).  Also row sums of |A|.
// Partial products of r = b - A x over column chunk ch = blockIdx.y:
// part[ch][i][q] = sum_j A(i, j) x(j, q) and rowabs[ch][i] = sum_j |A(i, j)|,
// summed in column order (deterministic).  The chunk's rows of x are staged
// in LDS ([j][q]); streaming as k_mx_rowmax.
template <int VEC>
__global__ void __launch_bounds__(128) k_mx_resid_part(const double *__restrict__ A, long lda, int n,
                                                       int cw, const double *__restrict__ x,
                                                       double *part, double *rowabs) {
  extern __shared__ double sx[];  // cw x NR
  const int ch = blockIdx.y;
  const int j0 = ch * cw, nj = min(n - j0, cw);
  for (int t = threadIdx.x; t < nj * NR; t += 128) {
    const int j = t / NR, q = t % NR;
    sx[t] = x[j0 + j + (long)n * q];
  }
  __syncthreads();
  const int i = blockIdx.x * SR + 2 * threadIdx.x;
  if (i >= n) return;
  const bool two = i + 1 < n;
  const double *Ap = A + i + lda * (long)j0;
  double s0[NR], s1[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) s0[q] = s1[q] = 0.0;
  double sa0 = 0.0, sa1 = 0.0;
  int j = 0;
  for (; j + SU <= nj; j += SU) {
    double2 a[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) a[u] = load_pair<VEC>(Ap + lda * (long)(j + u), two);
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      sa0 += fabs(a[u].x);
      sa1 += fabs(a[u].y);
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        const double xv = sx[(j + u) * NR + q];
        s0[q] = fma(a[u].x, xv, s0[q]);
        s1[q] = fma(a[u].y, xv, s1[q]);
      }
    }
  }
  for (; j < nj; ++j) {
    const double2 a = load_pair<0>(Ap + lda * (long)j, two);
    sa0 += fabs(a.x);
    sa1 += fabs(a.y);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const double xv = sx[j * NR + q];
      s0[q] = fma(a.x, xv, s0[q]);
      s1[q] = fma(a.y, xv, s1[q]);
    }
  }
  double *p = part + ((long)ch * n + i) * NR;
#pragma unroll
  for (int q = 0; q < NR; ++q) p[q] = s0[q];
  if (two)
#pragma unroll
    for (int q = 0; q < NR; ++q) p[NR + q] = s1[q];
  if (rowabs) {
    rowabs[(long)ch * n + i] = sa0;
    if (two) rowabs[(long)ch * n + i + 1] = sa1;
  }
}

__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = v > w ? v : w;
  }
  return v;
}

// r = b - sum of partials; norms[q] = max |r_q|, norms[8+q] = max |x_q|,
// norms[16] = max row sum of |A| (when rowabs is given).  Doubles are >= 0,
// so their bit patterns order like the values.
__global__ void __launch_bounds__(256) k_mx_resid_fin(int n, int nch, const double *__restrict__ part,
                                                      const double *__restrict__ b, double *r,
                                                      const double *__restrict__ x,
                                                      const double *__restrict__ rowabs,
                                                      unsigned long long *norms) {
  __shared__ unsigned long long sm[4][2 * NR + 1];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long v[2 * NR + 1];
#pragma unroll
  for (int q = 0; q < 2 * NR + 1; ++q) v[q] = 0ull;
  if (i < n) {
    double s[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) s[q] = b[i + (long)n * q];
    for (int ch = 0; ch < nch; ++ch) {
      const double *p = part + ((long)ch * n + i) * NR;
#pragma unroll
      for (int q = 0; q < NR; ++q) s[q] -= p[q];
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      r[i + (long)n * q] = s[q];
      v[q] = (unsigned long long)__double_as_longlong(fabs(s[q]));
      v[NR + q] = (unsigned long long)__double_as_longlong(fabs(x[i + (long)n * q]));
    }
    if (rowabs) {
      double ra = 0.0;
      for (int ch = 0; ch < nch; ++ch) ra += rowabs[(long)ch * n + i];
      v[2 * NR] = (unsigned long long)__double_as_longlong(ra);
    }
  }
#pragma unroll
  for (int q = 0; q < 2 * NR + 1; ++q) v[q] = wave_max_u64(v[q]);
  if (lane == 0)
    for (int q = 0; q < 2 * NR + 1; ++q) sm[wv][q] = v[q];
  __syncthreads();
  if (threadIdx.x < 2 * NR + 1) {
    unsigned long long m = sm[0][threadIdx.x];
    for (int w = 1; w < 4; ++w) m = m > sm[w][threadIdx.x] ? m : sm[w][threadIdx.x];
    if (m) atomicMax(&norms[threadIdx.x], m);
  }
}

__global__ void k_mx_zero_u64(unsigned long long *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Copy an n x nrhs column-major block into an n x 8 one (zero columns past nrhs).
__global__ void k_mx_pack_rhs(const double *__restrict__ B, long ldb, int n, int nrhs, double *o) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < NR; ++q) o[i + (long)n * q] = q < nrhs ? B[i + ldb * q] : 0.0;
}

__global__ void k_mx_unpack_rhs(const double *__restrict__ x, int n, int nrhs, double *B, long ldb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nrhs; ++q) B[i + ldb * q] = x[i + (long)n * q];
}

__global__ void k_mx_copy_a(const double *__restrict__ A, long lda, int n, double *o) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)n * n) return;
  long i = q % n, j = q / n;
  o[q] = A[i + lda * j];
}

// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------
struct MixedWs {
  int dev = -1;
  DevBuf<float> W;
  DevBuf<double> rs;
  DevBuf<unsigned> rowmax;
  DevBuf<char> Lf, Uf;
  DevBuf<int> piv, perm, npiv;
  DevBuf<float> dinv, Linv, Uinv;
  DevBuf<unsigned> flags;       // [0] max |l|, [1] zero pivot, [2] solve error, [3..] ready
  DevBuf<float> y;
  DevBuf<double> b8, x8, r8, part, rowabs, A64;
  DevBuf<unsigned long long> norms;
  DevBuf<int> ipiv64;
  unsigned gen = 0;
  // page-locked targets of the asynchronous read-backs
  struct Readback {
    unsigned flags[2];                    // max |l| bits, zero pivot
    unsigned long long norms[2 * NR + 1];  // residual and solution norms, ||A||
    unsigned err;                         // triangular-solve time-out
  };
  HostBuf<Readback> hb;
  // streams of the single-system solve (default priority): s1 runs the solve,
  // s2 the lookahead's trailing remainder on all CUs but a reserved few
  hipStream_t s1 = nullptr, s2 = nullptr;
  hipEvent_t ev_in = nullptr, ev_start = nullptr, ev_panel = nullptr, ev_rest = nullptr;
};
MixedWs g_ws;

// Compute units of the current device (queried once per device).
int mx_cus() {
  static int cached[64] = {};
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev >= 0 && dev < 64 && cached[dev] > 0) return cached[dev];
  hipDeviceProp_t pr;
  int n = 64;
  if (hipGetDeviceProperties(&pr, dev) == hipSuccess && pr.multiProcessorCount > 0)
    n = pr.multiProcessorCount;
  if (dev >= 0 && dev < 64) cached[dev] = n;
  return n;
}

size_t lds_diag() { return sizeof(double) * (size_t)NB * LDD; }
size_t lds_trsm() {
  size_t l = sizeof(float) * ((size_t)NB * LDX + (size_t)NB * LDT + 4 * SP * SP);
  size_t u = sizeof(float) * ((size_t)64 * LDT + (size_t)NB * LDT + 4 * SP * SP);
  return std::max(l, u);
}

#define MX_LAUNCH_CHECK(what)                                                     \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess) {                                                       \
      set_error("%s launch failed: %s", what, hipGetErrorString(e_));             \
      return DUST_E_HIP;                                                          \
    }                                                                             \
  } while (0)

// maxwg caps the workgroups (each then takes several tiles).
int mx_gemm(int f16, float *W, long ld, int r0, int r1, int c0, int c1, const void *Lf,
            const void *Uf, long pbytes, int P, hipStream_t st, int maxwg = 1 << 30) {
  if (r1 <= r0 || c1 <= c0 || P <= 0) return 0;
  const int nrt = (r1 - r0) / 128, ntiles = nrt * ((c1 - c0) / 128);
  const int grid = std::min(ntiles, maxwg);
  GemmItem g{W, ld, r0, c0, nrt, P, ntiles, Lf, Uf, pbytes};
  if (f16) k_mx_gemm<1, 0><<<grid, 256, 0, st>>>(g, nullptr, nullptr, 0);
  else k_mx_gemm<0, 0><<<grid, 256, 0, st>>>(g, nullptr, nullptr, 0);
  MX_LAUNCH_CHECK("update gemm");
  return 0;
}

// Two update regions of the same P block columns in one launch.
int mx_gemm2(int f16, float *W, long ld, int r0, int r1, int c0, int c1, int s0, int s1, int d0,
             int d1, const void *Lf, const void *Uf, long pbytes, int P, hipStream_t st) {
  if (P <= 0) return 0;
  auto item = [&](int ra, int rb, int ca, int cb) {
    const int nrt = rb > ra ? (rb - ra) / 128 : 0;
    const int nt = (rb > ra && cb > ca) ? nrt * ((cb - ca) / 128) : 0;
    return GemmItem{W, ld, ra, ca, std::max(nrt, 1), P, nt, Lf, Uf, pbytes};
  };
  const GemmItem a = item(r0, r1, c0, c1), b = item(s0, s1, d0, d1);
  const int grid = a.ntiles + b.ntiles;
  if (grid == 0) return 0;
  if (f16) k_mx_gemm2<1><<<grid, 256, 0, st>>>(a, b);
  else k_mx_gemm2<0><<<grid, 256, 0, st>>>(a, b);
  MX_LAUNCH_CHECK("update gemm");
  return 0;
}

// Factorisation of W (np x np).  Returns 0 or a HIP error; *bad is set when a
// block failed the pivot check or hit a zero pivot.
int mx_factor_launch(MixedWs &w, int np, int f16, hipStream_t st) {
  const int nblk = np / NB;
  const long ld = np;
  const size_t esz = f16 ? sizeof(_Float16) : sizeof(float);
  const long pbytes = (long)np * NB * (long)esz;  // bytes per fragment panel
  // Lookahead: the trailing update of super-panel i is split into the next
  // super-panel's block column and block row (on `st`, needed at once) and the
  // remainder (on the second stream `s2`), which runs while `st` factors the
  // next super-panel.  The two fragment sets alternate between super-panels.
  hipStream_t s2 = w.s2;
  k_mx_fill_u32<<<1, 64, 0, st>>>(w.flags.p, 3, 0u);
  DUST_HIP(hipEventRecord(w.ev_start, st));
  DUST_HIP(hipStreamWaitEvent(s2, w.ev_start, 0));
  bool pending = false;  // a remainder update is queued on s2
  int sp = 0;
  for (int ks = 0; ks < np; ks += DEPTH * NB, ++sp) {
    char *Lf = w.Lf.p + (size_t)(sp & 1) * DEPTH * pbytes;
    char *Uf = w.Uf.p + (size_t)(sp & 1) * DEPTH * pbytes;
    const int nsub = std::min(DEPTH, (np - ks) / NB);
    for (int a = 0; a < nsub; ++a) {
      const int k = ks + a * NB;
      if (a > 0) {
        // block column a and block row a, updated by the sub-panels before it
        // (one launch: the two regions write disjoint tiles)
        DUST_TRY(mx_gemm2(f16, w.W.p, ld, k, np, k, k + NB, k, k + NB, k + NB, np, Lf, Uf, pbytes,
                          a, st));
      }
      DiagArgs da{w.W.p, ld, k, w.piv.p, w.perm.p, w.npiv.p, w.dinv.p, w.flags.p};
      k_mx_diag<0><<<1, 256, lds_diag(), st>>>(da, nullptr);
      MX_LAUNCH_CHECK("diagonal block");
      LaswpItem li{w.W.p, ld, np, k, w.piv.p, w.npiv.p};
      k_mx_laswp<0><<<ceil_div(np, 256), 256, 0, st>>>(li, nullptr, nullptr, 0);
      MX_LAUNCH_CHECK("laswp");
      const int nt = (np - k - NB) / 64;
      if (nt > 0) {
        TrsmArgs ta{w.W.p, ld, k, np, nt, f16, 0, w.dinv.p, Lf + (size_t)a * pbytes,
                    Uf + (size_t)a * pbytes, w.flags.p, nullptr, nullptr};
        k_mx_trsm<0><<<2 * nt, 256, lds_trsm(), st>>>(ta, nullptr, nullptr, 0);
        MX_LAUNCH_CHECK("block triangular solve");
      }
    }
    const int t0 = ks + nsub * NB;
    if (t0 >= np) break;
    const int t1 = std::min(np, t0 + DEPTH * NB);
    DUST_HIP(hipEventRecord(w.ev_panel, st));
    if (pending) DUST_HIP(hipStreamWaitEvent(st, w.ev_rest, 0));
    // the next super-panel's block column and block row
    DUST_TRY(mx_gemm2(f16, w.W.p, ld, t0, np, t0, t1, t0, t1, t1, np, Lf, Uf, pbytes, nsub, st));
    // the remainder on s2
    if (t1 < np) {
      DUST_HIP(hipStreamWaitEvent(s2, w.ev_panel, 0));
      DUST_TRY(mx_gemm(f16, w.W.p, ld, t1, np, t1, np, Lf, Uf, pbytes, nsub, s2));
      DUST_HIP(hipEventRecord(w.ev_rest, s2));
      pending = true;
    } else {
      pending = false;
    }
  }
  if (pending) DUST_HIP(hipStreamWaitEvent(st, w.ev_rest, 0));
  // 128x128 inverses of every diagonal block for the refinement solves
  TrsmArgs ia{w.W.p, ld, 0, np, 2, f16, 1, w.dinv.p, nullptr, nullptr, w.flags.p, w.Linv.p, w.Uinv.p};
  k_mx_trsm<0><<<dim3(4, nblk), 256, lds_trsm(), st>>>(ia, nullptr, nullptr, 0);
  MX_LAUNCH_CHECK("block inverses");
  // max |l| and the zero-pivot flag, read after the caller's synchronisation
  DUST_HIP(hipMemcpyAsync(w.hb.p->flags, w.flags.p, sizeof(w.hb.p->flags), hipMemcpyDeviceToHost, st));
  return 0;
}

// After the stream has drained: did a block fail the multiplier check or hit a
// zero pivot?
bool mx_factor_bad(const MixedWs &w, double tau) {
  float lmax;
  std::memcpy(&lmax, &w.hb.p->flags[0], sizeof(float));
  return w.hb.p->flags[1] != 0 || !(lmax <= (float)tau);
}

// y <- U^-1 L^-1 y (y: np x 8 FP32, in place).
int mx_solve(MixedWs &w, int np, hipStream_t st, int cus) {
  const int nblk = np / NB;
  const int G = std::min(nblk, cus);
  for (int upper = 0; upper < 2; ++upper) {
    if (++w.gen == 0) ++w.gen;  // generation 0 is the cleared state
    BlockSolveArgs<float> sa{w.W.p, (long)np, np, nblk, upper ? w.Uinv.p : w.Linv.p, w.y.p,
                             w.flags.p + 3 + upper * nblk, w.gen, w.flags.p + 2};
    // G <= #CUs small blocks, all resident at once (ordinary launch; the flag
    // waits are bounded)
    if (upper) k_block_solve<float, 1, 256><<<G, 256, 0, st>>>(sa);
    else k_block_solve<float, 0, 256><<<G, 256, 0, st>>>(sa);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("triangular solve launch failed: %s", hipGetErrorString(e));
      return DUST_E_HIP;
    }
  }
  return 0;
}

// r = b - A x and the norms (read back asynchronously into the workspace,
// with the triangular-solve time-out flag).
// Kernels only: the norms go to `norms` (2 NR + 1 words, zeroed by the caller).
int mx_residual_kernels(MixedWs &w, const double *dA, long lda, int n, bool with_anorm,
                        unsigned long long *norms, hipStream_t st) {
  int nch;
  const int cw = stream_chunk(n, &nch);
  dim3 g(ceil_div(n, SR), nch);
  const size_t lds = sizeof(double) * cw * NR;
  double *ra = with_anorm ? w.rowabs.p : nullptr;
  if (stream_vec(dA, lda, n)) k_mx_resid_part<1><<<g, 128, lds, st>>>(dA, lda, n, cw, w.x8.p, w.part.p, ra);
  else k_mx_resid_part<0><<<g, 128, lds, st>>>(dA, lda, n, cw, w.x8.p, w.part.p, ra);
  k_mx_resid_fin<<<ceil_div(n, 256), 256, 0, st>>>(n, nch, w.part.p, w.b8.p, w.r8.p, w.x8.p,
                                                   with_anorm ? w.rowabs.p : nullptr, norms);
  MX_LAUNCH_CHECK("residual");
  return 0;
}

int mx_residual_launch(MixedWs &w, const double *dA, long lda, int n, bool with_anorm,
                       hipStream_t st) {
  k_mx_zero_u64<<<1, 64, 0, st>>>(w.norms.p, 2 * NR + 1);
  DUST_TRY(mx_residual_kernels(w, dA, lda, n, with_anorm, w.norms.p, st));
  DUST_HIP(hipMemcpyAsync(w.hb.p->norms, w.norms.p, sizeof(w.hb.p->norms), hipMemcpyDeviceToHost, st));
  DUST_HIP(hipMemcpyAsync(&w.hb.p->err, w.flags.p + 2, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  return 0;
}

// After the stream has drained: the largest backward error over the
// right-hand sides, max_q ||r_q||_inf / (||A||_inf ||x_q||_inf).
double mx_residual_berr(const MixedWs &w, int nrhs, bool with_anorm, double *anorm) {
  double nm[2 * NR + 1];
  std::memcpy(nm, w.hb.p->norms, sizeof(nm));
  if (with_anorm) *anorm = nm[2 * NR];
  double be = 0.0;
  for (int q = 0; q < nrhs; ++q) {
    double den = *anorm * nm[NR + q];
    double e = den > 0.0 ? nm[q] / den : (nm[q] > 0.0 ? INFINITY : 0.0);
    be = std::max(be, e);
  }
  return be;
}

int mx_reserve(MixedWs &w, int n, int np) {
  const size_t nblk = np / NB;
  DUST_TRY(w.W.reserve((size_t)np * np));
  DUST_TRY(w.rs.reserve(np));
  DUST_TRY(w.rowmax.reserve(np));
  DUST_TRY(w.Lf.reserve((size_t)2 * DEPTH * np * NB * sizeof(float)));
  DUST_TRY(w.Uf.reserve((size_t)2 * DEPTH * np * NB * sizeof(float)));
  DUST_TRY(w.piv.reserve(np));
  DUST_TRY(w.perm.reserve(np));
  DUST_TRY(w.npiv.reserve(nblk));
  DUST_TRY(w.dinv.reserve(nblk * 8 * SP * SP));
  DUST_TRY(w.Linv.reserve(nblk * NB * NB));
  DUST_TRY(w.Uinv.reserve(nblk * NB * NB));
  if (w.flags.n < 3 + 2 * nblk) {
    // ready flags start at generation 0 (never a live generation)
    w.flags.n = 0;
    DUST_TRY(w.flags.reserve(3 + 2 * nblk));
    DUST_HIP(hipMemset(w.flags.p, 0, sizeof(unsigned) * w.flags.n));
  }
  DUST_TRY(w.y.reserve((size_t)np * NR));
  DUST_TRY(w.b8.reserve((size_t)n * NR));
  DUST_TRY(w.x8.reserve((size_t)n * NR));
  DUST_TRY(w.r8.reserve((size_t)n * NR));
  int nch;
  (void)stream_chunk(n, &nch);
  DUST_TRY(w.part.reserve((size_t)nch * n * NR));
  DUST_TRY(w.rowabs.reserve((size_t)nch * n));
  DUST_TRY(w.norms.reserve(2 * NR + 1));
  DUST_TRY(w.hb.reserve(1));
  return 0;
}

// ---------------------------------------------------------------------------
// Drivers.  A system is factored once; its right-hand sides are refined in
// chunks of 8.  Every step is split into launches (ending in asynchronous
// read-backs into the workspace) and a host-side finish, so that a batch of
// systems can advance in lockstep with one synchronisation per step.
// ---------------------------------------------------------------------------
struct MixedRun {
  MixedWs *w = nullptr;
  hipStream_t st = nullptr;
  const double *dA = nullptr;
  long lda = 0;
  double *dB = nullptr;
  long ldb = 0;
  int n = 0, np = 0, nrhs = 0, prec = 0, max_iter = 0;
  double tol = 0.0;
  bool factored = false;  // tuned factors usable
  bool fp64 = false;      // FP64 partial-pivoting factors in w->A64 / w->ipiv64
  // current chunk
  int q0 = 0, nq = 0, nit = 0, slow = 0, it_report = 0;
  double anorm = 0.0, berr = INFINITY, prev = INFINITY, resid_report = 0.0;
  bool refining = false, chunk_done = false;
  int rc = 0;
};

int set_attributes() {
  static std::once_flag attr;
  hipError_t ae = hipSuccess;
  std::call_once(attr, [&] {
    const void *diag[] = {(const void *)k_mx_diag<0>, (const void *)k_mx_diag<1>};
    const void *trsm[] = {(const void *)k_mx_trsm<0>, (const void *)k_mx_trsm<1>};
    for (int b = 0; b < 2 && ae == hipSuccess; ++b) {
      ae = hipFuncSetAttribute(diag[b], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_diag());
      if (ae == hipSuccess)
        ae = hipFuncSetAttribute(trsm[b], hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds_trsm());
    }
  });
  DUST_HIP(ae);
  return 0;
}

// W = R A (launches only).
int run_convert_launch(MixedRun &r) {
  MixedWs &w = *r.w;
  const int f16 = r.prec == DUST_LU_FP16S;
  const int n = r.n, np = r.np;
  DUST_TRY(set_attributes());
  if (f16) {
    k_mx_fill_u32<<<ceil_div(np, 256), 256, 0, r.st>>>(w.rowmax.p, np, 0u);
    int nch;
    const int cw = stream_chunk(n, &nch);
    dim3 g(ceil_div(n, SR), nch);
    if (stream_vec(r.dA, r.lda, n)) k_mx_rowmax<1><<<g, 128, 0, r.st>>>(r.dA, r.lda, n, cw, w.rowmax.p);
    else k_mx_rowmax<0><<<g, 128, 0, r.st>>>(r.dA, r.lda, n, cw, w.rowmax.p);
  }
  k_mx_rowscale<<<ceil_div(np, 256), 256, 0, r.st>>>(w.rowmax.p, n, np, f16, w.rs.p);
  k_mx_convert<<<dim3(np / 256 + (np % 256 != 0), np / 4), 256, 0, r.st>>>(r.dA, r.lda, n, np,
                                                                         w.rs.p, w.W.p);
  MX_LAUNCH_CHECK("convert");
  return 0;
}

// W = R A and the tuned factorisation (launches only).
int run_factor_launch(MixedRun &r) {
  DUST_TRY(run_convert_launch(r));
  return mx_factor_launch(*r.w, r.np, r.prec == DUST_LU_FP16S, r.st);
}

// FP64 partial-pivoting factors (fallback and DUST_LU_FP64); synchronous.
int run_fp64_factor(MixedRun &r) {
  MixedWs &w = *r.w;
  DUST_TRY(w.A64.reserve((size_t)r.n * r.n));
  DUST_TRY(w.ipiv64.reserve(r.n));
  long tot = (long)r.n * r.n;
  k_mx_copy_a<<<(int)((tot + 255) / 256), 256, 0, r.st>>>(r.dA, r.lda, r.n, w.A64.p);
  MX_LAUNCH_CHECK("copy");
  int info = 0;
  DUST_TRY(lu_getrf_dev(r.n, r.n, w.A64.p, r.n, w.ipiv64.p, r.st, 64, &info));
  if (info) return info;
  r.fp64 = true;
  return 0;
}

// Start a chunk of right-hand sides: pack them into w.b8.
int run_chunk_begin(MixedRun &r, int q0) {
  r.q0 = q0;
  r.nq = std::min(NR, r.nrhs - q0);
  r.nit = 0;
  r.slow = 0;
  r.berr = r.prev = INFINITY;
  r.chunk_done = false;
  k_mx_pack_rhs<<<ceil_div(r.n, 256), 256, 0, r.st>>>(r.dB + r.ldb * q0, r.ldb, r.n, r.nq,
                                                      r.w->b8.p);
  MX_LAUNCH_CHECK("pack");
  r.refining = r.factored && !r.fp64;
  return 0;
}

// One refinement sweep (launches): d = U^-1 L^-1 P R r (r = b first), x += d,
// r = b - A x with its norms.  Split around the triangular solves so that a
// batch can run the solves of all its systems in one launch.
int run_sweep_gather(MixedRun &r) {
  MixedWs &w = *r.w;
  ++r.nit;
  k_mx_gather<<<ceil_div(r.np, 256), 256, 0, r.st>>>(r.nit == 1 ? w.b8.p : w.r8.p, r.n, r.n, r.np,
                                                     r.nq, w.perm.p, w.rs.p, w.y.p);
  MX_LAUNCH_CHECK("gather");
  return 0;
}

int run_sweep_update(MixedRun &r) {
  k_mx_update_x<<<ceil_div(r.n, 256), 256, 0, r.st>>>(r.w->x8.p, r.n, r.n, r.w->y.p, r.nit > 1);
  MX_LAUNCH_CHECK("update");
  return 0;
}

int run_sweep_launch(MixedRun &r) {
  DUST_TRY(run_sweep_gather(r));
  DUST_TRY(mx_solve(*r.w, r.np, r.st, mx_cus()));
  DUST_TRY(run_sweep_update(r));
  return mx_residual_launch(*r.w, r.dA, r.lda, r.n, r.nit == 1, r.st);
}

// After the stream has drained: convergence test of the sweep.
int run_sweep_finish(MixedRun &r) {
  MixedWs &w = *r.w;
  DUST_CHECK(w.hb.p->err == 0, DUST_E_HIP, "triangular solve flag wait timed out");
  r.berr = mx_residual_berr(w, r.nq, r.nit == 1, &r.anorm);
  if (r.berr <= r.tol) {
    r.refining = false;
    r.chunk_done = true;
  } else if (!(r.berr < 0.5 * r.prev)) {
    if (++r.slow >= 2) {
      r.refining = false;
      r.chunk_done = r.berr <= 1e-14;  // stalled at working accuracy
    }
  } else {
    r.slow = 0;
  }
  r.prev = r.berr;
  if (r.refining && r.nit >= std::max(1, r.max_iter)) r.refining = false;
  return 0;
}

// Finish a chunk: FP64 path when refinement did not converge, then unpack.
int run_chunk_end(MixedRun &r) {
  MixedWs &w = *r.w;
  int nit = r.nit;
  if (!r.chunk_done) {
    if (!r.fp64) {
      int info = run_fp64_factor(r);
      if (info > 0) return info;
      DUST_TRY(info);
    }
    DUST_HIP(hipMemcpyAsync(w.x8.p, w.b8.p, sizeof(double) * (size_t)r.n * NR,
                            hipMemcpyDeviceToDevice, r.st));
    DUST_TRY(lu_getrs_dev(r.n, NR, w.A64.p, r.n, w.ipiv64.p, w.x8.p, r.n, r.st));
    DUST_TRY(mx_residual_launch(w, r.dA, r.lda, r.n, true, r.st));
    DUST_HIP(hipStreamSynchronize(r.st));
    r.berr = mx_residual_berr(w, r.nq, true, &r.anorm);
    nit = r.prec == DUST_LU_FP64 ? 0 : -1;
  }
  k_mx_unpack_rhs<<<ceil_div(r.n, 256), 256, 0, r.st>>>(w.x8.p, r.n, r.nq, r.dB + r.ldb * r.q0,
                                                        r.ldb);
  MX_LAUNCH_CHECK("unpack");
  r.it_report = (nit < 0 || r.it_report < 0) ? -1 : std::max(r.it_report, nit);
  r.resid_report = std::max(r.resid_report, r.berr);
  return 0;
}

int run_init(MixedRun &r, MixedWs &w, int n, int nrhs, const double *dA, long lda, double *dB,
             long ldb, int prec, double tol, int max_iter, hipStream_t st) {
  DUST_CHECK(n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n), DUST_E_ARG,
             "invalid dgesv_mixed dimensions");
  DUST_CHECK(prec == DUST_LU_FP64 || prec == DUST_LU_FP32 || prec == DUST_LU_FP16S, DUST_E_ARG,
             "unknown precision %d", prec);
  r = MixedRun();
  r.w = &w;
  r.st = st;
  r.dA = dA;
  r.lda = lda;
  r.dB = dB;
  r.ldb = ldb;
  r.n = n;
  r.np = ((n + NB - 1) / NB) * NB;
  r.nrhs = nrhs;
  r.prec = prec;
  r.tol = tol;
  r.max_iter = max_iter;
  if (n == 0 || nrhs == 0) return 0;
  return mx_reserve(w, n, r.np);
}

// Streams of the single-system solve, created on first use.  The lookahead
// stream s2 is kept off the first g_reserved_cus compute units (a CU-masked
// stream; an ordinary one if masking is unavailable), so that the
// latency-bound diagonal-block kernel and the panel updates of the next step
// always find free CUs while the trailing remainder runs.  A CU-masked stream
// synchronises with the legacy default stream, so the solve itself runs on
// its own non-blocking stream s1, ordered after the caller's stream.
int g_reserved_cus = 32;  // tuning knob "mixed_reserved_cus"

int mx_streams(MixedWs &w) {
  if (w.s1) return 0;
  DUST_HIP(hipStreamCreateWithFlags(&w.s1, hipStreamNonBlocking));
  const int cus = mx_cus();
  std::vector<uint32_t> mask((cus + 31) / 32, 0u);
  for (int c = g_reserved_cus; c < cus; ++c) mask[c / 32] |= 1u << (c % 32);
  if (g_reserved_cus <= 0 || cus <= 2 * g_reserved_cus ||
      hipExtStreamCreateWithCUMask(&w.s2, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
    (void)hipGetLastError();
    w.s2 = nullptr;
    DUST_HIP(hipStreamCreateWithFlags(&w.s2, hipStreamNonBlocking));
  }
  DUST_HIP(hipEventCreateWithFlags(&w.ev_in, hipEventDisableTiming));
  DUST_HIP(hipEventCreateWithFlags(&w.ev_start, hipEventDisableTiming));
  DUST_HIP(hipEventCreateWithFlags(&w.ev_panel, hipEventDisableTiming));
  DUST_HIP(hipEventCreateWithFlags(&w.ev_rest, hipEventDisableTiming));
  return 0;
}

int mixed_dev(MixedWs &w, int n, int nrhs, const double *dA, long lda, double *dB, long ldb, int prec,
              double tol, int max_iter, int *iters, double *resid, hipStream_t cst) {
  if (iters) *iters = 0;
  if (resid) *resid = 0.0;
  DUST_TRY(mx_streams(w));
  // the solve runs on s1 after the work already queued on the caller's stream
  const hipStream_t st = w.s1;
  DUST_HIP(hipEventRecord(w.ev_in, cst));
  DUST_HIP(hipStreamWaitEvent(st, w.ev_in, 0));
  MixedRun r;
  DUST_TRY(run_init(r, w, n, nrhs, dA, lda, dB, ldb, prec, tol, max_iter, st));
  if (n == 0 || nrhs == 0) return 0;
  if (prec != DUST_LU_FP64) {
    DUST_TRY(run_factor_launch(r));
    DUST_HIP(hipStreamSynchronize(st));
    r.factored = !mx_factor_bad(w, TAU);
  }
  for (int q0 = 0; q0 < nrhs; q0 += NR) {
    DUST_TRY(run_chunk_begin(r, q0));
    while (r.refining) {
      DUST_TRY(run_sweep_launch(r));
      DUST_HIP(hipStreamSynchronize(st));
      DUST_TRY(run_sweep_finish(r));
    }
    int rc = run_chunk_end(r);
    if (rc) {
      (void)hipStreamSynchronize(st);  // leave no work of this call running
      return rc;
    }
  }
  DUST_HIP(hipStreamSynchronize(st));
  if (iters) *iters = r.it_report;
  if (resid) *resid = r.resid_report;
  return 0;
}

// ---------------------------------------------------------------------------
// Batches of independent systems of different sizes, on one stream.  The
// factorisation advances every system by one block column per step, and each
// step is one launch per kernel (panel update, diagonal blocks, interchanges,
// triangular solves, trailing update) whose workgroups are dealt out to the
// systems that have that block column: a region table with a prefix sum of
// workgroup counts, searched by each workgroup.  The refinement solves of all
// systems are one launch per triangle.  Launch count: ~5 per block column of
// the largest system instead of ~5 per block column of every system, and the
// latency-bound diagonal-block chains of all systems run side by side.
// ---------------------------------------------------------------------------
enum { BK_GEMM16, BK_GEMM32, BK_DIAG, BK_LASWP, BK_TRSM, BK_SOLVE_L, BK_SOLVE_U };

struct BatchLaunch {
  int kind, m, blocks;
  size_t items, starts;  // byte offsets in the plan
};

struct BatchPlan {
  std::vector<char> blob;
  std::vector<BatchLaunch> launches;
  size_t append(const void *p, size_t bytes) {
    size_t off = (blob.size() + 15) & ~(size_t)15;
    blob.resize(off + bytes);
    std::memcpy(blob.data() + off, p, bytes);
    return off;
  }
  // One launch over `items`; item i gets blocks[i] workgroups (items with
  // none are dropped).
  template <class T>
  void add(int kind, const std::vector<T> &items, const std::vector<int> &blocks) {
    std::vector<T> it;
    std::vector<int> st(1, 0);
    for (size_t i = 0; i < items.size(); ++i)
      if (blocks[i] > 0) {
        it.push_back(items[i]);
        st.push_back(st.back() + blocks[i]);
      }
    if (it.empty()) return;
    BatchLaunch L;
    L.kind = kind;
    L.m = (int)it.size();
    L.blocks = st.back();
    L.items = append(it.data(), it.size() * sizeof(T));
    L.starts = append(st.data(), st.size() * sizeof(int));
    launches.push_back(L);
  }
};

constexpr int NW = 2 * NR + 1;  // norm words per system

struct BatchState {
  std::vector<std::unique_ptr<MixedWs>> ws;
  hipStream_t st = nullptr;
  DevBuf<char> dplan;
  HostBuf<char> hplan;
  DevBuf<unsigned> flags;            // [3 i ..]: max |l|, zero pivot, solve time-out
  DevBuf<unsigned long long> norms;  // [NW i ..]
  HostBuf<unsigned> hflags;
  HostBuf<unsigned long long> hnorms;
};
BatchState g_batch;

// Upload the plan (page-locked staging; the previous plan's copy has been
// waited for by the caller's last synchronisation) and launch it.
int plan_run(BatchState &B, const BatchPlan &P, hipStream_t st) {
  if (P.launches.empty()) return 0;
  DUST_TRY(B.hplan.reserve(P.blob.size()));
  DUST_TRY(B.dplan.reserve(P.blob.size()));
  std::memcpy(B.hplan.p, P.blob.data(), P.blob.size());
  DUST_HIP(hipMemcpyAsync(B.dplan.p, B.hplan.p, P.blob.size(), hipMemcpyHostToDevice, st));
  const char *base = B.dplan.p;
  for (const BatchLaunch &L : P.launches) {
    const int *starts = (const int *)(base + L.starts);
    const void *items = base + L.items;
    switch (L.kind) {
      case BK_GEMM16:
        k_mx_gemm<1, 1><<<L.blocks, 256, 0, st>>>(GemmItem{}, (const GemmItem *)items, starts, L.m);
        break;
      case BK_GEMM32:
        k_mx_gemm<0, 1><<<L.blocks, 256, 0, st>>>(GemmItem{}, (const GemmItem *)items, starts, L.m);
        break;
      case BK_DIAG:  // one workgroup per item
        k_mx_diag<1><<<L.blocks, 256, lds_diag(), st>>>(DiagArgs{}, (const DiagArgs *)items);
        break;
      case BK_LASWP:
        k_mx_laswp<1><<<L.blocks, 256, 0, st>>>(LaswpItem{}, (const LaswpItem *)items, starts, L.m);
        break;
      case BK_TRSM:
        k_mx_trsm<1><<<L.blocks, 256, lds_trsm(), st>>>(TrsmArgs{}, (const TrsmArgs *)items, starts,
                                                       L.m);
        break;
      case BK_SOLVE_L:
        k_block_solve_batch<float, 0, 256><<<L.blocks, 256, 0, st>>>(
            (const BlockSolveArgs<float> *)items, starts, L.m);
        break;
      case BK_SOLVE_U:
        k_block_solve_batch<float, 1, 256><<<L.blocks, 256, 0, st>>>(
            (const BlockSolveArgs<float> *)items, starts, L.m);
        break;
    }
    MX_LAUNCH_CHECK("batched launch");
  }
  return 0;
}

// Factorisation of every system of the batch (W already converted): the
// steps of mx_factor_launch without the lookahead, one fragment set.
void batch_factor_plan(std::vector<MixedRun> &runs, BatchState &B, int f16, BatchPlan &P) {
  int maxblk = 0;
  for (const MixedRun &r : runs) maxblk = std::max(maxblk, r.np / NB);
  const int gk = f16 ? BK_GEMM16 : BK_GEMM32;
  const size_t esz = f16 ? sizeof(_Float16) : sizeof(float);
  for (int kk = 0; kk < maxblk; ++kk) {
    const int k = kk * NB, a = kk % DEPTH;
    std::vector<GemmItem> pg, tg;
    std::vector<DiagArgs> dg;
    std::vector<LaswpItem> lg;
    std::vector<TrsmArgs> sg;
    std::vector<int> pb, tb, db, lb, sb;
    for (size_t i = 0; i < runs.size(); ++i) {
      MixedRun &r = runs[i];
      MixedWs &w = *r.w;
      const int np = r.np;
      if (r.n == 0 || kk >= np / NB) continue;
      const long ld = np;
      const long pbytes = (long)np * NB * (long)esz;
      char *Lf = w.Lf.p, *Uf = w.Uf.p;
      unsigned *fl = B.flags.p + 3 * i;
      if (a > 0) {
        // block column and block row of this step, updated by the sub-panels
        // before it in the super-panel
        pg.push_back(GemmItem{w.W.p, ld, k, k, (np - k) / NB, a, (np - k) / NB, Lf, Uf, pbytes});
        pb.push_back((np - k) / NB);
        pg.push_back(GemmItem{w.W.p, ld, k, k + NB, 1, a, (np - k - NB) / NB, Lf, Uf, pbytes});
        pb.push_back((np - k - NB) / NB);
      }
      dg.push_back(DiagArgs{w.W.p, ld, k, w.piv.p, w.perm.p, w.npiv.p, w.dinv.p, fl});
      db.push_back(1);
      lg.push_back(LaswpItem{w.W.p, ld, np, k, w.piv.p, w.npiv.p});
      lb.push_back(ceil_div(np, 256));
      const int nt = (np - k - NB) / 64;
      sg.push_back(TrsmArgs{w.W.p, ld, k, np, nt, f16, 0, w.dinv.p, Lf + (size_t)a * pbytes,
                            Uf + (size_t)a * pbytes, fl, nullptr, nullptr});
      sb.push_back(2 * nt);
      if (a == DEPTH - 1 && k + NB < np) {
        const int t0 = k + NB, nrt = (np - t0) / NB;
        tg.push_back(GemmItem{w.W.p, ld, t0, t0, nrt, DEPTH, nrt * nrt, Lf, Uf, pbytes});
        tb.push_back(nrt * nrt);
      }
    }
    P.add(gk, pg, pb);
    P.add(BK_DIAG, dg, db);
    P.add(BK_LASWP, lg, lb);
    P.add(BK_TRSM, sg, sb);
    P.add(gk, tg, tb);
  }
  std::vector<TrsmArgs> ig;
  std::vector<int> ib;
  for (size_t i = 0; i < runs.size(); ++i) {
    MixedRun &r = runs[i];
    MixedWs &w = *r.w;
    ig.push_back(TrsmArgs{w.W.p, (long)r.np, 0, r.np, 2, f16, 1, w.dinv.p, nullptr, nullptr,
                          B.flags.p + 3 * i, w.Linv.p, w.Uinv.p});
    ib.push_back(r.n > 0 ? 4 * (r.np / NB) : 0);
  }
  P.add(BK_TRSM, ig, ib);
}

// Both triangular solves of the systems in `act`, in one plan.  Workgroups
// per system in proportion to its block rows, about one per CU in total.
void batch_solve_plan(std::vector<MixedRun> &runs, const std::vector<int> &act, BatchState &B,
                      int cus, BatchPlan &P) {
  long tot = 0;
  for (int i : act) tot += runs[i].np / NB;
  const long budget = std::max<long>(cus, (long)act.size());
  for (int upper = 0; upper < 2; ++upper) {
    std::vector<BlockSolveArgs<float>> items;
    std::vector<int> blocks;
    for (int i : act) {
      MixedWs &w = *runs[i].w;
      const int np = runs[i].np, nblk = np / NB;
      if (++w.gen == 0) ++w.gen;  // generation 0 is the cleared state
      items.push_back(BlockSolveArgs<float>{w.W.p, (long)np, np, nblk,
                                            upper ? w.Uinv.p : w.Linv.p, w.y.p,
                                            w.flags.p + 3 + upper * nblk, w.gen,
                                            B.flags.p + 3 * i + 2});
      blocks.push_back((int)std::max(1L, std::min<long>(nblk, budget * nblk / std::max(tot, 1L))));
    }
    P.add(upper ? BK_SOLVE_U : BK_SOLVE_L, items, blocks);
  }
}

int mixed_batch(int nb, const int *n, const double *const *dA, const int *lda, double *const *dB,
                const int *ldb, int nrhs, int prec, double tol, int max_iter, int *iters,
                double *resid) {
  BatchState &B = g_batch;
  if (!B.st) DUST_HIP(hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking));
  while ((int)B.ws.size() < nb) B.ws.emplace_back(new MixedWs());
  hipStream_t st = B.st;
  std::vector<MixedRun> runs(nb);
  for (int i = 0; i < nb; ++i)
    DUST_TRY(run_init(runs[i], *B.ws[i], n[i], nrhs, dA[i], lda[i], dB[i], ldb[i], prec, tol,
                      max_iter, st));
  DUST_TRY(B.flags.reserve((size_t)3 * nb));
  DUST_TRY(B.norms.reserve((size_t)NW * nb));
  DUST_TRY(B.hflags.reserve((size_t)3 * nb));
  DUST_TRY(B.hnorms.reserve((size_t)NW * nb));
  DUST_TRY(set_attributes());
  const int cus = mx_cus();
  if (nrhs == 0) return 0;
  k_mx_fill_u32<<<ceil_div(3 * nb, 256), 256, 0, st>>>(B.flags.p, 3 * nb, 0u);
  MX_LAUNCH_CHECK("flags");
  if (prec != DUST_LU_FP64) {
    const int f16 = prec == DUST_LU_FP16S;
    for (int i = 0; i < nb; ++i)
      if (runs[i].n > 0) DUST_TRY(run_convert_launch(runs[i]));
    BatchPlan P;
    batch_factor_plan(runs, B, f16, P);
    DUST_TRY(plan_run(B, P, st));
    DUST_HIP(hipMemcpyAsync(B.hflags.p, B.flags.p, sizeof(unsigned) * 3 * nb,
                            hipMemcpyDeviceToHost, st));
    DUST_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nb; ++i) {
      if (runs[i].n == 0) continue;  // no workspace
      std::memcpy(runs[i].w->hb.p->flags, B.hflags.p + 3 * i, sizeof(unsigned) * 2);
      runs[i].factored = !mx_factor_bad(*runs[i].w, TAU);
    }
  }
  for (int q0 = 0; q0 < nrhs; q0 += NR) {
    for (int i = 0; i < nb; ++i)
      if (runs[i].n > 0) DUST_TRY(run_chunk_begin(runs[i], q0));
    while (true) {
      std::vector<int> act;
      for (int i = 0; i < nb; ++i)
        if (runs[i].n > 0 && runs[i].refining) act.push_back(i);
      if (act.empty()) break;
      k_mx_zero_u64<<<ceil_div(NW * nb, 256), 256, 0, st>>>(B.norms.p, NW * nb);
      for (int i : act) DUST_TRY(run_sweep_gather(runs[i]));
      BatchPlan P;
      batch_solve_plan(runs, act, B, cus, P);
      DUST_TRY(plan_run(B, P, st));
      for (int i : act) {
        MixedRun &r = runs[i];
        DUST_TRY(run_sweep_update(r));
        DUST_TRY(mx_residual_kernels(*r.w, r.dA, r.lda, r.n, r.nit == 1, B.norms.p + (size_t)NW * i,
                                     st));
      }
      DUST_HIP(hipMemcpyAsync(B.hnorms.p, B.norms.p, sizeof(unsigned long long) * NW * nb,
                              hipMemcpyDeviceToHost, st));
      DUST_HIP(hipMemcpyAsync(B.hflags.p, B.flags.p, sizeof(unsigned) * 3 * nb,
                              hipMemcpyDeviceToHost, st));
      DUST_HIP(hipStreamSynchronize(st));
      for (int i : act) {
        MixedWs &w = *runs[i].w;
        std::memcpy(w.hb.p->norms, B.hnorms.p + (size_t)NW * i, sizeof(w.hb.p->norms));
        w.hb.p->err = B.hflags.p[3 * i + 2];
        DUST_TRY(run_sweep_finish(runs[i]));
      }
    }
    for (int i = 0; i < nb; ++i)
      if (runs[i].n > 0) {
        int rc = run_chunk_end(runs[i]);
        if (rc) return rc;
      }
  }
  DUST_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < nb; ++i) {
    if (iters) iters[i] = runs[i].it_report;
    if (resid) resid[i] = runs[i].resid_report;
  }
  return 0;
}

void ws_release(MixedWs &w) {
  w.W.release(); w.rs.release(); w.rowmax.release(); w.Lf.release(); w.Uf.release();
  w.piv.release(); w.perm.release(); w.npiv.release(); w.dinv.release(); w.Linv.release();
  w.Uinv.release(); w.flags.release(); w.y.release(); w.b8.release(); w.x8.release();
  w.r8.release(); w.part.release(); w.rowabs.release(); w.A64.release(); w.norms.release();
  w.ipiv64.release(); w.hb.release();
  w.gen = 0;
  w.dev = -1;
  for (hipEvent_t *e : {&w.ev_in, &w.ev_start, &w.ev_panel, &w.ev_rest})
    if (*e) {
      (void)hipEventDestroy(*e);
      *e = nullptr;
    }
  for (hipStream_t *q : {&w.s1, &w.s2})
    if (*q) {
      (void)hipStreamDestroy(*q);
      *q = nullptr;
    }
}

}  // namespace

void mixed_release() {
  ws_release(g_ws);
  BatchState &B = g_batch;
  for (auto &w : B.ws) ws_release(*w);
  B.ws.clear();
  if (B.st) (void)hipStreamDestroy(B.st);
  B.st = nullptr;
  B.dplan.release();
  B.hplan.release();
  B.flags.release();
  B.norms.release();
  B.hflags.release();
  B.hnorms.release();
}

int mixed_reserved_cus() { return g_reserved_cus; }

// The workspaces (and their CU-masked streams) are rebuilt on the next solve.
int mixed_set_reserved_cus(int v) {
  DUST_HIP(hipDeviceSynchronize());
  mixed_release();
  g_reserved_cus = v;
  return 0;
}

}  // namespace dust

using namespace dust;

extern "C" {

int dust_dgesv_mixed_dev(int n, int nrhs, const double *dA, int lda, double *dB, int ldb, int prec,
                         double tol, int max_iter, int *iters, double *resid, void *stream) {
  return mixed_dev(g_ws, n, nrhs, dA, lda, dB, ldb, prec, tol, max_iter, iters, resid,
                   (hipStream_t)stream);
}

int dust_dgesv_mixed(int n, int nrhs, const double *A, int lda, double *B, int ldb, int prec,
                     double tol, int max_iter, int *iters, double *resid) {
  DUST_CHECK(A && B && n >= 0 && nrhs >= 0 && lda >= std::max(1, n) && ldb >= std::max(1, n),
             DUST_E_ARG, "invalid dgesv_mixed arguments");
  if (n == 0 || nrhs == 0) {
    if (iters) *iters = 0;
    if (resid) *resid = 0.0;
    return 0;
  }
  double *dA = nullptr, *dB = nullptr;
  DUST_HIP(hipMalloc(&dA, sizeof(double) * (size_t)n * n));
  if (hipMalloc(&dB, sizeof(double) * (size_t)n * nrhs) != hipSuccess) {
    (void)hipFree(dA);
    set_error("hipMalloc failed");
    return DUST_E_NOMEM;
  }
  int rc = 0;
  if (hipMemcpy2D(dA, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                  hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy2D(dB, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                  hipMemcpyHostToDevice) != hipSuccess) {
    set_error("upload failed");
    rc = DUST_E_HIP;
  }
  if (!rc) rc = mixed_dev(g_ws, n, nrhs, dA, n, dB, n, prec, tol, max_iter, iters, resid, nullptr);
  if (!rc && hipMemcpy2D(B, sizeof(double) * ldb, dB, sizeof(double) * n, sizeof(double) * n, nrhs,
                         hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("download failed");
    rc = DUST_E_HIP;
  }
  (void)hipFree(dA);
  (void)hipFree(dB);
  return rc;
}

// Independent systems of different sizes on one stream: the factorisations
// advance together one block column per step (one launch per kernel and step
// for the whole batch), then refinement sweeps of all unfinished systems with
// both triangular solves of the batch in one launch each and one
// synchronisation per sweep.  Device pointers.
int dust_dgesv_mixed_batch_dev(int nb, const int *n, const double *const *dA, const int *lda,
                               double *const *dB, const int *ldb, int nrhs, int prec, double tol,
                               int max_iter, int *iters, double *resid) {
  DUST_CHECK(nb >= 0 && (nb == 0 || (n && dA && lda && dB && ldb)), DUST_E_ARG,
             "invalid batch arguments");
  DUST_CHECK(nrhs >= 0, DUST_E_ARG, "invalid batch arguments");
  if (iters)
    for (int i = 0; i < nb; ++i) iters[i] = 0;
  if (resid)
    for (int i = 0; i < nb; ++i) resid[i] = 0.0;
  if (nb == 0) return 0;
  return mixed_batch(nb, n, dA, lda, dB, ldb, nrhs, prec, tol, max_iter, iters, resid);
}

}  // extern "C"
