// Assembly of DUST's linear system for a body with a panel wake, on gfx950:
// initialize_linsys and assemble_linsys of src/linsys/mod_linsys.f90
// (109-272, 279-344) with the row routines of the elements.
//
// Reference behaviour:
//   build_row_static_surfpan / build_row_surfpan (src/aeroel/mod_surfpan.f90:
//     448-481, 377-439), build_row_static_vortlatt / build_row_vortlatt
//     (src/aeroel/mod_vortlatt.f90:201-237, 143-192): one AIC entry and one
//     right-hand-side coefficient b1 per (row, column) pair from compute_pot
//     (surface-panel rows) or compute_psi (vortex-lattice rows) of the
//     column element (mod_surfpan.f90:579-632, mod_vortlatt.f90:331-391);
//   add_wake_surfpan / add_wake_vortlatt (mod_surfpan.f90:489-535,
//     mod_vortlatt.f90:280-322): the implicit first wake row adds +-a into
//     A(ie, ind1/ind2); the explicit wake panels subtract a*mag from b(ie);
//   add_expl_surfpan / add_expl_vortlatt (mod_surfpan.f90:543-570,
//     mod_vortlatt.f90:245-272) with compute_pot/psi_liftlin
//     (src/aeroel/mod_liftlin.f90:154-206), the static explicit elements
//     through L_static.
//
// Layout: B(ldb, rank) keeps b_static in its static block (written once by
// the initialisation) and this step's b1 coefficients in the other blocks, so
// the right-hand side is one product b = B s with s_j = nor_j.(ub_j - u_inf -
// uvort_j) (ub only for moving columns, as the reference's two loops).  The
// AIC kernels are one column per blockIdx.y with rows across threads
// (coalesced column writes); the wake and explicit corrections are one thread
// per row walking the wake list in the reference's order, so every A(ie, ind)
// receives its additions in the reference's sequence.
#include <algorithm>
#include <vector>

#include "aic_internal.h"
#include "common.h"
#include "lu_internal.h"

namespace dust {
namespace {

struct AsmParams {
  double ffr, rran, rcut;
};

// compute_pot / compute_psi of element g (kind ck) seen by the row element
// of kind rk at position pos with normal nor, for the pair (i, j).
__device__ inline void pair_entry(int rk, const double pos[3], const double nor[3], int ck,
                                  const Panel &g, int i, int j, const AsmParams &q, double &a,
                                  double &b1) {
  if (rk == DUST_ROW_SURFPAN) {
    const double dou = (i != j) ? pot_doublet(g, pos, q.ffr) : -2.0 * kPi;
    a = -dou;
    if (ck == DUST_ROW_SURFPAN) {
      int w = 0;
      b1 = pot_source(g, dou, pos, &w);
    } else {
      b1 = 0.0;
    }
  } else {
    double v[3];
    vel_doublet(g, pos, q.ffr, q.rran, q.rcut, v);
    a = dot3(v, nor);
    if (ck == DUST_ROW_SURFPAN) {
      double vs[3];
      vel_source(g, pos, q.ffr, vs);
      b1 = -dot3(vs, nor);
    } else {
      b1 = (i == j) ? 4.0 * kPi : 0.0;
    }
  }
}

// A(i, j) and B(i, j) for rows [r0, r1) x columns [c0, c1).
__global__ void __launch_bounds__(256)
    k_asm_block(int r0, int r1, int c0, const int *__restrict__ kind, const Panel *__restrict__ P,
                AsmParams q, double *A, long lda, double *B, long ldb) {
  const int i = r0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = c0 + blockIdx.y;
  if (i >= r1) return;
  const Panel &gj = P[j];
  double a, b1;
  pair_entry(kind[i], P[i].cen, P[i].nor, kind[j], gj, i, j, q, a, b1);
  A[i + lda * j] = a;
  B[i + ldb * j] = b1;
}

// L_static(ie, j) of the static rows and the static explicit elements.
__global__ void __launch_bounds__(256)
    k_asm_lstatic(int nst, const int *__restrict__ kind, const Panel *__restrict__ P,
                  const Panel *__restrict__ Pe, AsmParams q, double *L, long ldl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nst) return;
  double a, b1;
  pair_entry(kind[i], P[i].cen, P[i].nor, DUST_ROW_VORTLATT, Pe[j], 1, 2, q, a, b1);
  L[i + ldl * j] = a;
}

// s_j = nor_j . (ub_j - u_inf - uvort_j); ub only for the moving columns.
__global__ void k_asm_sigma(int n, int nst, const Panel *__restrict__ P,
                            const double *__restrict__ ub, const double *__restrict__ uvort,
                            double u0, double u1, double u2, double *s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double uinf[3] = {u0, u1, u2};
  double w[3];
  for (int e = 0; e < 3; ++e)
    w[e] = ((ub && j >= nst) ? ub[3 * j + e] : 0.0) - uinf[e] - (uvort ? uvort[3 * j + e] : 0.0);
  s[j] = dot3(P[j].nor, w);
}

struct WakeRowArgs {
  int r0, r1, nst, rank;
  int wc0_static, wc1_static;  // wake column range of the static rows
  int n_wake, n_impl;
  int n_expl, ec0_static;      // explicit range [ec0, n_expl) of the static rows
  const int *kind;
  const Panel *P, *Pw, *Pe;
  const double *wmag, *emag;
  const int *impl_ind;         // (2, n_impl) 1-based, 0 = none
  const double *L;             // L_static (ldl, nstatic_expl) or NULL
  long ldl;
  AsmParams q;
  double *A;
  long lda;
  double *b;                   // NULL: the right-hand side is not formed
};

// add_wake then add_expl for the rows [r0, r1): static rows (ie < nst) take
// the wake columns [wc0_static, wc1_static) and the explicit elements from
// ec0_static on (the others through L_static), the other rows everything.
__global__ void __launch_bounds__(256) k_asm_wake_rows(WakeRowArgs W) {
  const int ie = W.r0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (ie >= W.r1) return;
  const bool st = ie < W.nst;
  const int c0 = st ? W.wc0_static : 0, c1 = st ? W.wc1_static : W.rank;
  const int rk = W.kind[ie];
  const Panel &gi = W.P[ie];
  for (int k = 0; k < W.n_impl; ++k) {
    const int i1 = W.impl_ind[2 * k] - 1, i2 = W.impl_ind[2 * k + 1] - 1;
    const bool in1 = i1 >= c0 && i1 < c1, in2 = i2 >= c0 && i2 < c1;
    // a surface-panel row takes a wake panel only when both generating
    // elements are in range (a missing lower element, ind2 = 0, never is)
    if (rk == DUST_ROW_SURFPAN ? (in1 && in2) : in1) {
      double a, bb;
      pair_entry(rk, gi.cen, gi.nor, DUST_ROW_VORTLATT, W.Pw[k], 1, 2, W.q, a, bb);
      W.A[ie + W.lda * i1] += a;
      if (i2 >= 0) W.A[ie + W.lda * i2] -= a;
    }
  }
  if (!W.b) return;
  double acc = W.b[ie];
  for (int k = W.n_impl; k < W.n_wake; ++k) {
    double a, bb;
    pair_entry(rk, gi.cen, gi.nor, DUST_ROW_VORTLATT, W.Pw[k], 1, 2, W.q, a, bb);
    acc -= a * W.wmag[k];
  }
  const int e0 = st ? W.ec0_static : 0;
  for (int j = 0; j < e0; ++j) acc -= W.L[ie + W.ldl * j] * W.emag[j];
  for (int j = e0; j < W.n_expl; ++j) {
    double a, bb;
    pair_entry(rk, gi.cen, gi.nor, DUST_ROW_VORTLATT, W.Pe[j], 1, 2, W.q, a, bb);
    acc -= a * W.emag[j];
  }
  W.b[ie] = acc;
}

// dst(r, c) = src(ri[r], ci[c]) for r < nr, c < nc (the pressure-system copies
// of A by idSurfPan).
__global__ void k_asm_gather(int nr, int nc, const int *__restrict__ ri, const int *__restrict__ ci,
                             const double *__restrict__ src, long lds, double *dst, long ldd) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= nr) return;
  dst[r + ldd * c] = src[ri[r] + lds * ci[c]];
}

// Bernoulli right-hand side of build_row_surfpan (mod_surfpan.f90:405-437):
// b_pres(p) = sum_q B(id_sp(p), id_sp(q)) bern(id_sp(q)).
__global__ void k_asm_bpres(int nsp, const int *__restrict__ id, const double *__restrict__ B,
                            long ldb, const double *__restrict__ bern, double *bp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nsp) return;
  const long row = id[p];
  double acc = 0.0;
  for (int q = 0; q < nsp; ++q) {
    const int j = id[q];
    acc += B[row + ldb * j] * bern[j];
  }
  bp[p] = acc;
}

struct AsmWs {
  DevBuf<Panel> P, Pw, Pe;
  DevBuf<double> s, part;
};
AsmWs g_asm;

int check_desc(const dust_linsys_desc *d) {
  DUST_CHECK(d, DUST_E_ARG, "null system description");
  DUST_CHECK(d->rank >= 1 && d->nstatic >= 0 && d->nstatic <= d->rank, DUST_E_ARG,
             "invalid rank / nstatic");
  DUST_CHECK(d->kind && d->n_ver && d->ver, DUST_E_ARG, "missing element arrays");
  DUST_CHECK(d->n_sp >= 0 && d->nstatic_sp >= 0 && d->nstatic_sp <= d->n_sp &&
                 d->n_sp <= d->rank && (d->n_sp == 0 || d->id_sp),
             DUST_E_ARG, "invalid surface-panel list");
  DUST_CHECK(d->n_wake >= 0 && d->n_impl >= 0 && d->n_impl <= d->n_wake &&
                 (d->n_wake == 0 || (d->wake_nver && d->wake_ver)) &&
                 (d->n_impl == 0 || d->impl_ind) &&
                 (d->n_wake == d->n_impl || d->wake_mag),
             DUST_E_ARG, "invalid wake-panel list");
  DUST_CHECK(d->n_expl >= 0 && d->nstatic_expl >= 0 && d->nstatic_expl <= d->n_expl &&
                 (d->n_expl == 0 || (d->expl_nver && d->expl_ver && d->expl_mag)),
             DUST_E_ARG, "invalid explicit-element list");
  return 0;
}

int asm_geometry(const dust_linsys_desc *d, hipStream_t st) {
  AsmWs &w = g_asm;
  DUST_TRY(w.P.reserve(d->rank));
  k_geo<<<ceil_div(d->rank, 128), 128, 0, st>>>(d->rank, d->n_ver, d->ver, w.P.p);
  if (d->n_wake > 0) {
    DUST_TRY(w.Pw.reserve(d->n_wake));
    k_geo<<<ceil_div(d->n_wake, 128), 128, 0, st>>>(d->n_wake, d->wake_nver, d->wake_ver,
                                                    w.Pw.p);
  }
  if (d->n_expl > 0) {
    DUST_TRY(w.Pe.reserve(d->n_expl));
    k_geo<<<ceil_div(d->n_expl, 128), 128, 0, st>>>(d->n_expl, d->expl_nver, d->expl_ver,
                                                    w.Pe.p);
  }
  DUST_HIP(hipGetLastError());
  return 0;
}

void asm_block(const dust_linsys_desc *d, int r0, int r1, int c0, int c1, const AsmParams &q,
               double *A, long lda, double *B, long ldb, hipStream_t st) {
  if (r1 <= r0 || c1 <= c0) return;
  dim3 grid(ceil_div(r1 - r0, 256), c1 - c0);
  k_asm_block<<<grid, 256, 0, st>>>(r0, r1, c0, d->kind, g_asm.P.p, q, A, lda, B, ldb);
}

WakeRowArgs wake_args(const dust_linsys_desc *d, const AsmParams &q, double *A, long lda,
                      const double *L, long ldl, double *b) {
  WakeRowArgs W{};
  W.r0 = 0;
  W.r1 = d->rank;
  W.nst = d->nstatic;
  W.rank = d->rank;
  W.n_wake = d->n_wake;
  W.n_impl = d->n_impl;
  W.n_expl = d->n_expl;
  W.ec0_static = d->nstatic_expl;
  W.kind = d->kind;
  W.P = g_asm.P.p;
  W.Pw = g_asm.Pw.p;
  W.Pe = g_asm.Pe.p;
  W.wmag = d->wake_mag;
  W.emag = d->expl_mag;
  W.impl_ind = d->impl_ind;
  W.L = L;
  W.ldl = ldl;
  W.q = q;
  W.A = A;
  W.lda = lda;
  W.b = b;
  return W;
}

}  // namespace

void asm_release() {
  g_asm.P.release();
  g_asm.Pw.release();
  g_asm.Pe.release();
  g_asm.s.release();
  g_asm.part.release();
}

}  // namespace dust

using namespace dust;

extern "C" {

int dust_initialize_linsys_dev(const dust_linsys_desc *d, double *A, int lda, double *B, int ldb,
                               double *L_static, int ldl, double *A_pres, int ldp, int *ipiv,
                               int *ipiv_pres, void *stream) {
  DUST_TRY(check_desc(d));
  const int n = d->rank, nst = d->nstatic, nsp = d->n_sp, nss = d->nstatic_sp;
  DUST_CHECK(A && B && lda >= n && ldb >= n, DUST_E_ARG, "invalid A / B");
  DUST_CHECK(d->nstatic_expl == 0 || (L_static && ldl >= std::max(1, nst)), DUST_E_ARG,
             "L_static missing or too small");
  DUST_CHECK(nsp == 0 || (A_pres && ldp >= nsp), DUST_E_ARG, "A_pres missing or too small");
  DUST_CHECK(nst == 0 || ipiv, DUST_E_ARG, "ipiv missing");
  DUST_CHECK(nss == 0 || ipiv_pres, DUST_E_ARG, "ipiv_pres missing");
  hipStream_t st = (hipStream_t)stream;
  const AsmParams q{d->ff_ratio_doublet, d->rankine_rad, d->cutoff_rad};
  DUST_TRY(asm_geometry(d, st));
  // build_row_static of every static row (174-179)
  asm_block(d, 0, nst, 0, nst, q, A, lda, B, ldb, st);
  if (nst > 0 && d->nstatic_expl > 0) {
    dim3 grid(ceil_div(nst, 256), d->nstatic_expl);
    k_asm_lstatic<<<grid, 256, 0, st>>>(nst, d->kind, g_asm.P.p, g_asm.Pe.p, q, L_static, ldl);
  }
  // A_pres copied before the wake correction (184-188); its static block is
  // the part defined now, the rest is refreshed by every assembly
  if (nss > 0) {
    dim3 grid(ceil_div(nss, 256), nss);
    k_asm_gather<<<grid, 256, 0, st>>>(nss, nss, d->id_sp, d->id_sp, A, lda, A_pres, ldp);
  }
  // static wake correction (191-198): implicit panels on static columns
  if (nst > 0 && d->n_impl > 0) {
    WakeRowArgs W = wake_args(d, q, A, lda, L_static, ldl, nullptr);
    W.r1 = nst;
    W.wc0_static = 0;
    W.wc1_static = nst;
    k_asm_wake_rows<<<ceil_div(nst, 256), 256, 0, st>>>(W);
  }
  DUST_HIP(hipGetLastError());
  // one-time LU of the two static blocks (236, 255)
  int info = 0;
  if (nst > 0) DUST_TRY(lu_getrf_dev(nst, nst, A, lda, ipiv, st, 64, &info));
  if (info) return info;
  if (nss > 0) DUST_TRY(lu_getrf_dev(nss, nss, A_pres, ldp, ipiv_pres, st, 64, &info));
  return info;
}

int dust_assemble_linsys_dev(const dust_linsys_desc *d, double *A, int lda, double *B, int ldb,
                             const double *L_static, int ldl, double *b, double *A_pres, int ldp,
                             double *b_pres, void *stream) {
  DUST_TRY(check_desc(d));
  const int n = d->rank, nst = d->nstatic, nsp = d->n_sp, nss = d->nstatic_sp;
  DUST_CHECK(A && B && b && lda >= n && ldb >= n, DUST_E_ARG, "invalid A / B / b");
  DUST_CHECK(d->nstatic_expl == 0 || nst == 0 || (L_static && ldl >= nst), DUST_E_ARG,
             "L_static missing or too small");
  DUST_CHECK(!A_pres || ldp >= nsp, DUST_E_ARG, "A_pres leading dimension too small");
  DUST_CHECK(!b_pres || (d->bern && nsp > 0), DUST_E_ARG, "b_pres needs bern and surface panels");
  hipStream_t st = (hipStream_t)stream;
  const AsmParams q{d->ff_ratio_doublet, d->rankine_rad, d->cutoff_rad};
  DUST_TRY(asm_geometry(d, st));
  // (2) rows: static rows over the moving columns, moving rows over all
  asm_block(d, 0, nst, nst, n, q, A, lda, B, ldb, st);
  asm_block(d, nst, n, 0, n, q, A, lda, B, ldb, st);
  // right-hand side b = B s
  AsmWs &w = g_asm;
  DUST_TRY(w.s.reserve(n));
  DUST_TRY(w.part.reserve((size_t)rhs_split(n) * n));
  k_asm_sigma<<<ceil_div(n, 256), 256, 0, st>>>(n, nst, w.P.p, d->ub, d->uvort, d->u_inf[0],
                                                d->u_inf[1], d->u_inf[2], w.s.p);
  rhs_gemv(n, 1, B, ldb, w.s.p, w.part.p, b, n, st);
  if (b_pres) k_asm_bpres<<<ceil_div(nsp, 256), 256, 0, st>>>(nsp, d->id_sp, B, ldb, d->bern,
                                                               b_pres);
  // (3) pressure-matrix copies of the moving blocks
  if (A_pres && nsp > nss) {
    if (nss > 0) {
      dim3 g1(ceil_div(nss, 256), nsp - nss);
      k_asm_gather<<<g1, 256, 0, st>>>(nss, nsp - nss, d->id_sp, d->id_sp + nss, A, lda,
                                       A_pres + (long)ldp * nss, ldp);
    }
    dim3 g2(ceil_div(nsp - nss, 256), nsp);
    k_asm_gather<<<g2, 256, 0, st>>>(nsp - nss, nsp, d->id_sp + nss, d->id_sp, A, lda,
                                     A_pres + nss, ldp);
  }
  // (4) wake and explicit corrections
  WakeRowArgs W = wake_args(d, q, A, lda, L_static, ldl, b);
  W.wc0_static = nst;
  W.wc1_static = n;
  k_asm_wake_rows<<<ceil_div(n, 256), 256, 0, st>>>(W);
  DUST_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Host-pointer entry points: the description's arrays, the matrices and the
// vectors are staged through device memory; synchronous.
// ---------------------------------------------------------------------------
namespace dust {
namespace {

struct Stage {
  std::vector<void *> bufs;
  hipError_t err = hipSuccess;
  Stage() = default;
  Stage(const Stage &) = delete;
  ~Stage() {
    for (void *p : bufs) (void)hipFree(p);
  }
  template <typename T>
  T *alloc(size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
      err = hipErrorOutOfMemory;
      return nullptr;
    }
    bufs.push_back(p);
    return (T *)p;
  }
  template <typename T>
  T *put(const T *h, size_t count) {
    if (!h) return nullptr;
    T *d = alloc<T>(count);
    if (d && hipMemcpy(d, h, sizeof(T) * count, hipMemcpyHostToDevice) != hipSuccess)
      err = hipErrorInvalidValue;
    return d;
  }
  template <typename T>
  int get(T *h, const T *d, size_t count) {
    if (!h || !d || count == 0) return 0;
    DUST_HIP(hipMemcpy(h, d, sizeof(T) * count, hipMemcpyDeviceToHost));
    return 0;
  }
};

// Device copy of a host description.
int stage_desc(Stage &S, const dust_linsys_desc *h, dust_linsys_desc *d) {
  DUST_TRY(check_desc(h));
  const size_t n = h->rank;
  for (size_t p = 0; p < n; ++p)
    DUST_CHECK(h->n_ver[p] == 3 || h->n_ver[p] == 4, DUST_E_ARG, "element %zu has %d vertices", p,
               h->n_ver[p]);
  for (int p = 0; p < h->n_sp; ++p)
    DUST_CHECK(h->id_sp[p] >= 0 && h->id_sp[p] < h->rank, DUST_E_ARG,
               "surface panel %d: element index %d out of range", p, h->id_sp[p]);
  for (int k = 0; k < h->n_impl; ++k)
    DUST_CHECK(h->impl_ind[2 * k] >= 1 && h->impl_ind[2 * k] <= h->rank &&
                   h->impl_ind[2 * k + 1] >= 0 && h->impl_ind[2 * k + 1] <= h->rank,
               DUST_E_ARG, "implicit wake panel %d: generating element out of range", k);
  for (int k = 0; k < h->n_wake; ++k)
    DUST_CHECK(h->wake_nver[k] == 3 || h->wake_nver[k] == 4, DUST_E_ARG,
               "wake panel %d has %d vertices", k, h->wake_nver[k]);
  for (int k = 0; k < h->n_expl; ++k)
    DUST_CHECK(h->expl_nver[k] == 3 || h->expl_nver[k] == 4, DUST_E_ARG,
               "explicit element %d has %d vertices", k, h->expl_nver[k]);
  *d = *h;
  d->kind = S.put(h->kind, n);
  d->n_ver = S.put(h->n_ver, n);
  d->ver = S.put(h->ver, 12 * n);
  d->ub = S.put(h->ub, 3 * n);
  d->uvort = S.put(h->uvort, 3 * n);
  d->bern = S.put(h->bern, n);
  d->id_sp = S.put(h->id_sp, (size_t)h->n_sp);
  d->wake_nver = S.put(h->wake_nver, (size_t)h->n_wake);
  d->wake_ver = S.put(h->wake_ver, 12 * (size_t)h->n_wake);
  d->wake_mag = S.put(h->wake_mag, (size_t)h->n_wake);
  d->impl_ind = S.put(h->impl_ind, 2 * (size_t)h->n_impl);
  d->expl_nver = S.put(h->expl_nver, (size_t)h->n_expl);
  d->expl_ver = S.put(h->expl_ver, 12 * (size_t)h->n_expl);
  d->expl_mag = S.put(h->expl_mag, (size_t)h->n_expl);
  DUST_HIP(S.err);
  return 0;
}

// Column-major (ld, ncol) host matrix to a packed device copy and back.
template <typename T>
T *put_mat(Stage &S, const T *h, int ld, int nrow, int ncol) {
  if (!h) return nullptr;
  T *d = S.template alloc<T>((size_t)nrow * ncol);
  if (d && hipMemcpy2D(d, sizeof(T) * nrow, h, sizeof(T) * ld, sizeof(T) * nrow, ncol,
                       hipMemcpyHostToDevice) != hipSuccess)
    S.err = hipErrorInvalidValue;
  return d;
}
int get_mat(double *h, int ld, const double *d, int nrow, int ncol) {
  if (!h || !d || nrow == 0 || ncol == 0) return 0;
  DUST_HIP(hipMemcpy2D(h, sizeof(double) * ld, d, sizeof(double) * nrow, sizeof(double) * nrow,
                       ncol, hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace
}  // namespace dust

extern "C" {

int dust_initialize_linsys(const dust_linsys_desc *desc, double *A, int lda, double *B, int ldb,
                           double *L_static, int ldl, double *A_pres, int ldp, int *ipiv,
                           int *ipiv_pres) {
  Stage S;
  dust_linsys_desc d;
  DUST_TRY(stage_desc(S, desc, &d));
  const int n = d.rank, nst = d.nstatic, nsp = d.n_sp, nse = d.nstatic_expl;
  DUST_CHECK(A && B && lda >= n && ldb >= n, DUST_E_ARG, "invalid A / B");
  DUST_CHECK(nse == 0 || (L_static && ldl >= std::max(1, nst)), DUST_E_ARG,
             "L_static missing or too small");
  DUST_CHECK(nsp == 0 || (A_pres && ldp >= nsp), DUST_E_ARG, "A_pres missing or too small");
  double *dA = put_mat(S, A, lda, n, n), *dB = put_mat(S, B, ldb, n, n);
  double *dL = nse ? put_mat(S, L_static, ldl, std::max(1, nst), nse) : nullptr;
  double *dAp = nsp ? put_mat(S, A_pres, ldp, nsp, nsp) : nullptr;
  int *dp = S.alloc<int>(std::max(1, nst)), *dpp = S.alloc<int>(std::max(1, d.nstatic_sp));
  DUST_HIP(S.err);
  int info = dust_initialize_linsys_dev(&d, dA, n, dB, n, dL, std::max(1, nst), dAp,
                                        std::max(1, nsp), dp, dpp, nullptr);
  if (info < 0) return info;
  DUST_HIP(hipDeviceSynchronize());
  DUST_TRY(get_mat(A, lda, dA, n, n));
  DUST_TRY(get_mat(B, ldb, dB, n, n));
  DUST_TRY(get_mat(L_static, ldl, dL, nst, nse));
  DUST_TRY(get_mat(A_pres, ldp, dAp, nsp, nsp));
  DUST_TRY(S.get(ipiv, dp, (size_t)nst));
  DUST_TRY(S.get(ipiv_pres, dpp, (size_t)d.nstatic_sp));
  return info;
}

int dust_assemble_linsys(const dust_linsys_desc *desc, double *A, int lda, double *B, int ldb,
                         const double *L_static, int ldl, double *b, double *A_pres, int ldp,
                         double *b_pres) {
  Stage S;
  dust_linsys_desc d;
  DUST_TRY(stage_desc(S, desc, &d));
  const int n = d.rank, nst = d.nstatic, nsp = d.n_sp, nse = d.nstatic_expl;
  DUST_CHECK(A && B && b && lda >= n && ldb >= n, DUST_E_ARG, "invalid A / B / b");
  DUST_CHECK(nse == 0 || nst == 0 || (L_static && ldl >= nst), DUST_E_ARG,
             "L_static missing or too small");
  DUST_CHECK(!A_pres || ldp >= nsp, DUST_E_ARG, "A_pres leading dimension too small");
  double *dA = put_mat(S, A, lda, n, n), *dB = put_mat(S, B, ldb, n, n);
  const double *dL = (nse && nst) ? put_mat(S, L_static, ldl, nst, nse) : nullptr;
  double *dAp = (A_pres && nsp) ? put_mat(S, A_pres, ldp, nsp, nsp) : nullptr;
  double *db = S.alloc<double>(n);
  double *dbp = b_pres ? S.alloc<double>(std::max(1, nsp)) : nullptr;
  DUST_HIP(S.err);
  DUST_TRY(dust_assemble_linsys_dev(&d, dA, n, dB, n, dL, std::max(1, nst), db, dAp,
                                    std::max(1, nsp), dbp, nullptr));
  DUST_HIP(hipDeviceSynchronize());
  DUST_TRY(get_mat(A, lda, dA, n, n));
  DUST_TRY(get_mat(B, ldb, dB, n, n));
  DUST_TRY(S.get(b, db, (size_t)n));
  DUST_TRY(get_mat(A_pres, ldp, dAp, nsp, nsp));
  DUST_TRY(S.get(b_pres, dbp, (size_t)nsp));
  return 0;
}

}  // extern "C"
