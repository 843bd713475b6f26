/*
 * linsys_oracle.c -- CPU restatement of the assembly of DUST's linear system
 * for a body with a panel wake: the static rows and the per-step moving rows,
 * the implicit first wake row, the explicit wake panels and explicit elements
 * of the right-hand side, the pressure-system copies, and the two drivers
 * initialize_linsys / assemble_linsys.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product library never does.
 *
 * Reference files (paths relative to the DUST source tree):
 *   src/aeroel/mod_surfpan.f90   build_row_surfpan 377-439,
 *                                build_row_static_surfpan 448-481,
 *                                add_wake_surfpan 489-535,
 *                                add_expl_surfpan 543-570,
 *                                compute_pot_surfpan 579-601,
 *                                compute_psi_surfpan 609-632
 *   src/aeroel/mod_vortlatt.f90  build_row_vortlatt 143-192,
 *                                build_row_static_vortlatt 201-237,
 *                                add_expl_vortlatt 245-272,
 *                                add_wake_vortlatt 280-322,
 *                                compute_pot_vortlatt 331-355,
 *                                compute_psi_vortlatt 365-391
 *   src/aeroel/mod_liftlin.f90   compute_pot_liftlin 154-174,
 *                                compute_psi_liftlin 183-206 (explicit
 *                                elements: the same doublet formulas)
 *   src/linsys/mod_linsys.f90    initialize_linsys 109-272,
 *                                assemble_linsys 279-344
 *
 * Indices are 0-based; column ranges are half-open [c0, c1) where the
 * reference passes 1-based inclusive (ista, iend).  The wind is the uniform
 * free stream (variable_wind without gusts).
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "aic_oracle.h"

void ao_vel_source(const double *g, const double pos[3], double ff_ratio_dou, double v[3]);

enum { AO_KIND_SURFPAN = 0, AO_KIND_VORTLATT = 1 };

/* System description (host arrays). */
typedef struct {
  int rank, nstatic;
  const int *kind;     /* (rank) AO_KIND_* */
  const double *G;     /* (G_SIZE, rank) geometry records */
  const double *ub;    /* (3, rank) body velocity at the centres, or NULL */
  const double *uvort; /* (3, rank) or NULL */
  const double *bern;  /* (rank) bernoulli_source of the surface panels, or NULL */
  double uinf[3];
  int n_sp, nstatic_sp;
  const int *id_sp;    /* (n_sp) implicit element of each surface panel */
  int n_wake, n_impl;
  const double *Gw;    /* (G_SIZE, n_wake) wake panels, the n_impl implicit ones first */
  const double *wmag;  /* (n_wake) */
  const int *impl_ind; /* (2, n_impl) pan_gen_elems_id: 1-based, 0 = none */
  int n_expl, nstatic_expl;
  const double *Ge;    /* (G_SIZE, n_expl) explicit elements, static first */
  const double *emag;  /* (n_expl) */
  double ff_ratio, rankine, cutoff;
} ao_linsys;

static const double *gofs(const double *G, int j) { return &G[(size_t)G_SIZE * j]; }

/* compute_pot of element g (kind) at pos for the pair (i, j). */
static void compute_pot(int kind, const double *g, const double *pos, int i, int j, double ffr,
                        double *A, double *b) {
  double dou = (i != j) ? ao_pot_doublet(g, pos, ffr) : -2.0 * AO_PI;
  *A = -dou;
  if (kind == AO_KIND_SURFPAN)
    *b = ao_pot_source(g, dou, pos, NULL);
  else
    *b = 0.0;
}

/* compute_psi of element g (kind) at pos with normal nor for the pair (i, j). */
static void compute_psi(int kind, const double *g, const double *pos, const double *nor, int i,
                        int j, const ao_linsys *L, double *A, double *b) {
  double v[3];
  ao_vel_doublet(g, pos, L->ff_ratio, L->rankine, L->cutoff, v);
  *A = ao_dot(v, nor);
  if (kind == AO_KIND_SURFPAN) {
    double vs[3];
    ao_vel_source(g, pos, L->ff_ratio, vs);
    *b = -ao_dot(vs, nor);
  } else {
    *b = (i == j) ? 4.0 * AO_PI : 0.0;
  }
}

/* The row element ie seen by element g of kind `kind` (pair indices i, j). */
static void row_entry(const ao_linsys *L, int ie, int kind, const double *g, int i, int j,
                      double *A, double *b) {
  const double *gi = gofs(L->G, ie);
  if (L->kind[ie] == AO_KIND_SURFPAN)
    compute_pot(kind, g, &gi[G_CEN], i, j, L->ff_ratio, A, b);
  else
    compute_psi(kind, g, &gi[G_CEN], &gi[G_NOR], i, j, L, A, b);
}

/* nor_j . (ub_j - wind - uvort_j) */
static double bc_coef(const ao_linsys *L, int j, int with_ub) {
  const double *nor = &gofs(L->G, j)[G_NOR];
  double w[3];
  for (int e = 0; e < 3; ++e)
    w[e] = (with_ub && L->ub ? L->ub[3 * j + e] : 0.0) - L->uinf[e] -
           (L->uvort ? L->uvort[3 * j + e] : 0.0);
  return ao_dot(nor, w);
}

/* build_row_static_{surfpan,vortlatt}: A(ie, c0:c1) and b_static(ie, c0:c1),
 * then L_static(ie, :) over the static explicit elements. */
void ao_build_row_static(const ao_linsys *L, int ie, int c0, int c1, double *A, int lda,
                         double *Bst, int ldb, double *Lst, int ldl) {
  for (int j = c0; j < c1; ++j) {
    double a, b1;
    row_entry(L, ie, L->kind[j], gofs(L->G, j), ie, j, &a, &b1);
    A[ie + (size_t)lda * j] = a;
    Bst[ie + (size_t)ldb * j] = b1;
  }
  for (int j = 0; j < L->nstatic_expl; ++j) {
    double a, b1;
    row_entry(L, ie, AO_KIND_VORTLATT, gofs(L->Ge, j), 1, 2, &a, &b1);
    if (Lst) Lst[ie + (size_t)ldl * j] = a;
  }
}

/* build_row_{surfpan,vortlatt}: b(ie) from the stored static coefficients of
 * the columns before c0 and from new entries of the columns [c0, c1), which
 * are also stored into A; the Bernoulli right-hand side b_pres for surface
 * panels (ipres = local index of ie among the surface panels, or -1). */
void ao_build_row(const ao_linsys *L, int ie, int c0, int c1, double *A, int lda,
                  const double *Bst, int ldb, const double *Bst_pres, int ldbp, int ipres,
                  double *b, double *b_pres) {
  double acc = 0.0;
  for (int j = 0; j < c0; ++j) acc += Bst[ie + (size_t)ldb * j] * bc_coef(L, j, 0);
  double accp = 0.0;
  if (ipres >= 0 && L->bern) {
    int m = c0 < L->nstatic_sp ? c0 : L->nstatic_sp;
    for (int j = 0; j < m; ++j)
      accp += Bst_pres[ipres + (size_t)ldbp * j] * L->bern[L->id_sp[j]];
  }
  for (int j = c0; j < c1; ++j) {
    double a, b1;
    row_entry(L, ie, L->kind[j], gofs(L->G, j), ie, j, &a, &b1);
    A[ie + (size_t)lda * j] = a;
    acc += b1 * bc_coef(L, j, 1);
    if (ipres >= 0 && L->bern && L->kind[j] == AO_KIND_SURFPAN) accp += b1 * L->bern[j];
  }
  b[ie] = acc;
  if (ipres >= 0 && b_pres) b_pres[ipres] = accp;
}

/* add_wake_{surfpan,vortlatt}: the implicit first-row wake panels add +-a
 * into A(ie, ind1/ind2) when their generating elements are in [c0, c1); the
 * explicit wake panels subtract a * mag from b(ie). */
void ao_add_wake(const ao_linsys *L, int ie, int c0, int c1, double *A, int lda, double *b) {
  const int surf = L->kind[ie] == AO_KIND_SURFPAN;
  for (int k = 0; k < L->n_impl; ++k) {
    const int i1 = L->impl_ind[2 * k] - 1, i2 = L->impl_ind[2 * k + 1] - 1;
    const int in1 = i1 >= c0 && i1 < c1, in2 = i2 >= c0 && i2 < c1;
    if (surf ? (in1 && in2) : in1) {
      double a, bb;
      row_entry(L, ie, AO_KIND_VORTLATT, gofs(L->Gw, k), 1, 2, &a, &bb);
      A[ie + (size_t)lda * i1] += a;
      if (i2 >= 0) A[ie + (size_t)lda * i2] -= a;
    }
  }
  for (int k = L->n_impl; k < L->n_wake; ++k) {
    double a, bb;
    row_entry(L, ie, AO_KIND_VORTLATT, gofs(L->Gw, k), 1, 2, &a, &bb);
    b[ie] -= a * L->wmag[k];
  }
}

/* add_expl_{surfpan,vortlatt}: explicit elements before e0 through the stored
 * L_static, those in [e0, e1) computed now. */
void ao_add_expl(const ao_linsys *L, int ie, int e0, int e1, const double *Lst, int ldl,
                 double *b) {
  for (int j = 0; j < e0; ++j) b[ie] -= Lst[ie + (size_t)ldl * j] * L->emag[j];
  for (int j = e0; j < e1; ++j) {
    double a, bb;
    row_entry(L, ie, AO_KIND_VORTLATT, gofs(L->Ge, j), 1, 2, &a, &bb);
    b[ie] -= a * L->emag[j];
  }
}

int ao_dgetrf(int m, int n, double *A, int lda, int *ipiv);

/* initialize_linsys (mod_linsys.f90:109-272): static rows, the pressure
 * matrix copied before the wake correction, the static wake correction,
 * b_static_pres, and the LU factors of the two static blocks.  A (lda, rank),
 * Bst (nst, nst), Lst (nst, nstatic_expl), A_pres / Bst_pres (n_sp, n_sp and
 * nstatic_sp, nstatic_sp); b (rank) is scratch here as in the reference.
 * Returns the first nonzero dgetrf info (static block, then pressure). */
int ao_initialize_linsys(const ao_linsys *L, double *A, int lda, double *Bst, double *Lst,
                         double *A_pres, double *Bst_pres, int *ipiv, int *ipiv_pres,
                         double *b) {
  const int nst = L->nstatic, nsp = L->n_sp, nss = L->nstatic_sp;
#pragma omp parallel for schedule(dynamic)
  for (int ie = 0; ie < nst; ++ie)
    ao_build_row_static(L, ie, 0, nst, A, lda, Bst, nst, Lst, nst);
  /* the pressure matrix copies A before the wake correction; only the
   * static block is defined at this point (the rest is refreshed by
   * assemble_linsys before every use) */
  for (int i = 0; i < nss; ++i)
    for (int j = 0; j < nss; ++j)
      A_pres[i + (size_t)nsp * j] = A[L->id_sp[i] + (size_t)lda * L->id_sp[j]];
#pragma omp parallel for
  for (int ie = 0; ie < nst; ++ie) ao_add_wake(L, ie, 0, nst, A, lda, b);
  for (int i = 0; i < nss; ++i)
    for (int j = 0; j < nss; ++j)
      Bst_pres[i + (size_t)nss * j] = Bst[L->id_sp[i] + (size_t)nst * L->id_sp[j]];
  int info = 0;
  if (nst > 0) info = ao_dgetrf(nst, nst, A, lda, ipiv);
  if (info) return info;
  if (nss > 0) info = ao_dgetrf(nss, nss, A_pres, nsp, ipiv_pres);
  return info;
}

/* assemble_linsys (mod_linsys.f90:279-344) after get_vort_vel (uvort is an
 * input): rows, the pressure-matrix copies of the moving blocks, then the
 * wake and explicit corrections.  b_pres (n_sp) when bern is given. */
void ao_assemble_linsys(const ao_linsys *L, double *A, int lda, const double *Bst,
                        const double *Lst, double *b, double *A_pres, const double *Bst_pres,
                        double *b_pres) {
  const int nst = L->nstatic, n = L->rank, nsp = L->n_sp, nss = L->nstatic_sp;
  int *g2l = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) g2l[i] = -1;
  for (int p = 0; p < nsp; ++p) g2l[L->id_sp[p]] = p;
#pragma omp parallel for schedule(dynamic)
  for (int ie = 0; ie < n; ++ie) {
    const int ip = L->kind[ie] == AO_KIND_SURFPAN ? g2l[ie] : -1;
    if (ie < nst)
      ao_build_row(L, ie, nst, n, A, lda, Bst, nst, Bst_pres, nss, ip, b, b_pres);
    else
      ao_build_row(L, ie, 0, n, A, lda, Bst, nst, Bst_pres, nss, ip, b, b_pres);
  }
  if (A_pres) {
    for (int i = 0; i < nss; ++i)
      for (int j = nss; j < nsp; ++j)
        A_pres[i + (size_t)nsp * j] = A[L->id_sp[i] + (size_t)lda * L->id_sp[j]];
    for (int i = nss; i < nsp; ++i)
      for (int j = 0; j < nsp; ++j)
        A_pres[i + (size_t)nsp * j] = A[L->id_sp[i] + (size_t)lda * L->id_sp[j]];
  }
#pragma omp parallel for schedule(dynamic)
  for (int ie = 0; ie < n; ++ie) {
    if (ie < nst) {
      ao_add_wake(L, ie, nst, n, A, lda, b);
      ao_add_expl(L, ie, L->nstatic_expl, L->n_expl, Lst, nst, b);
    } else {
      ao_add_wake(L, ie, 0, n, A, lda, b);
      ao_add_expl(L, ie, 0, L->n_expl, Lst, nst, b);
    }
  }
  free(g2l);
}
