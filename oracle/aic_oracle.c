/*
 * aic_oracle.c -- CPU restatement of DUST's panel influence kernels, the dense
 * AIC assembly and the LU solve of src/linsys.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product library never does.
 *
 * Reference files (paths relative to the DUST source tree):
 *   src/aeroel/mod_surfpan.f90   calc_geo_data_surfpan 890-957,
 *                                potential_calc_sou_surfpan 173-258,
 *                                velocity_calc_sou_surfpan 268-357,
 *                                compute_pot_surfpan 579-601,
 *                                build_row_static_surfpan 448-481
 *   src/aeroel/mod_doublet.f90   potential_calc_doublet 113-168,
 *                                velocity_calc_doublet 178-248
 *   src/aeroel/mod_vortlatt.f90  calc_geo_data_vortlatt 601-665,
 *                                compute_psi_vortlatt 365-391
 *   src/linsys/mod_linsys.f90    dgetrf/dgetrs call sites 236, 435, 473,
 *                                block path 362-465
 *   src/general/mod_param.f90    prev/next vertex tables 94-97
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "aic_oracle.h"

/* calc_geo_data_surfpan (mod_surfpan.f90:890-957); vortex rings use the same
 * fields (mod_vortlatt.f90:601-665) plus a guard on a zero normal. */
void ao_geo(int n, const int *nver, const double *ver /* (3,4,n) */, double *G /* (74,n) */) {
  for (int p = 0; p < n; ++p) {
    double *g = &G[(size_t)G_SIZE * p];
    const double *v = &ver[12 * p];
    int nv = nver[p];
    memset(g, 0, sizeof(double) * G_SIZE);
    g[G_NVER] = nv;
    memcpy(&g[G_VER], v, sizeof(double) * 3 * nv);
    for (int e = 0; e < 3; ++e) {
      double s = 0.0;
      for (int i = 0; i < nv; ++i) s += v[3 * i + e];
      g[G_CEN + e] = s / (double)nv;
    }
    double nor[3], a[3], b[3];
    if (nv == 4) {
      for (int e = 0; e < 3; ++e) { a[e] = v[6 + e] - v[e]; b[e] = v[9 + e] - v[3 + e]; }
    } else {
      for (int e = 0; e < 3; ++e) { a[e] = v[6 + e] - v[3 + e]; b[e] = v[e] - v[3 + e]; }
    }
    ao_cross(a, b, nor);
    double nn = ao_norm(nor);
    g[G_AREA] = 0.5 * nn;
    for (int e = 0; e < 3; ++e) g[G_NOR + e] = nor[e] / nn;
    double tanl[3];
    for (int e = 0; e < 3; ++e) tanl[e] = 0.5 * (v[3 * (nv - 1) + e] + v[e]) - g[G_CEN + e];
    double tn = ao_norm(tanl);
    for (int e = 0; e < 3; ++e) g[G_TANG + e] = tanl[e] / tn;
    ao_cross(&g[G_NOR], &g[G_TANG], &g[G_TANG + 3]);
    for (int is = 0; is < nv; ++is) {
      int nx = (nv == 3) ? next_tri[is] : next_qua[is];
      for (int e = 0; e < 3; ++e) g[G_EVEC + 3 * is + e] = v[3 * nx + e] - v[3 * is + e];
    }
    for (int is = 0; is < nv; ++is) g[G_ELEN + is] = ao_norm(&g[G_EVEC + 3 * is]);
    for (int is = 0; is < nv; ++is)
      for (int e = 0; e < 3; ++e) g[G_EUNI + 3 * is + e] = g[G_EVEC + 3 * is + e] / g[G_ELEN + is];
    for (int is = 0; is < nv; ++is) {
      g[G_COST + is] = ao_dot(&g[G_EUNI + 3 * is], &g[G_TANG]);
      g[G_SINT + is] = ao_dot(&g[G_EUNI + 3 * is], &g[G_TANG + 3]);
      double d[3];
      for (int e = 0; e < 3; ++e) d[e] = v[3 * is + e] - g[G_CEN + e];
      double z = ao_dot(d, &g[G_NOR]);
      for (int e = 0; e < 3; ++e) g[G_VERP + 3 * is + e] = v[3 * is + e] - g[G_NOR + e] * z;
    }
  }
}

/* potential_calc_doublet, mod_doublet.f90:113-168 */
double ao_pot_doublet(const double *g, const double pos[3], double ff_ratio) {
  int nv = (int)g[G_NVER];
  double rv[3] = {pos[0] - g[G_CEN], pos[1] - g[G_CEN + 1], pos[2] - g[G_CEN + 2]};
  double radius = ao_norm(rv);
  double zQ = ao_dot(rv, &g[G_NOR]);
  double dou;
  if (radius > ff_ratio * maxedge(g)) {
    dou = zQ * g[G_AREA] / (radius * radius * radius);
  } else {
    dou = 0.0;
    for (int i1 = 0; i1 < nv; ++i1) {
      int im1 = (nv == 3) ? prev_tri[i1] : prev_qua[i1];
      double ei[3];
      for (int e = 0; e < 3; ++e) ei[e] = -(pos[e] - g[G_VER + 3 * i1 + e]);
      double en = ao_norm(ei);
      for (int e = 0; e < 3; ++e) ei[e] /= en;
      const double *ev = &g[G_EVEC + 3 * i1], *evm = &g[G_EVEC + 3 * im1];
      double d1 = ao_dot(ei, ev), d2 = ao_dot(ei, evm);
      double ap1[3], am1[3];
      for (int e = 0; e < 3; ++e) {
        ap1[e] = ev[e] - ei[e] * d1;
        am1[e] = -evm[e] + ei[e] * d2;
      }
      double ap1n = ao_norm(ap1), am1n = ao_norm(am1);
      double cr[3];
      ao_cross(am1, ap1, cr);
      double sinB = ao_dot(ei, cr) / (ap1n * am1n);
      double cosB = ao_dot(am1, ap1) / (ap1n * am1n);
      dou += atan2(sinB, cosB);
    }
    double lim = (double)(nv - 2) * AO_PI;
    if (dou < -lim + 1.0e-5)
      dou = dou + lim;
    else if (dou > lim - 1.0e-5)
      dou = dou - lim;
  }
  return dou;
}

/* potential_calc_sou_surfpan, mod_surfpan.f90:173-258.  The far-field ratio
 * is the routine's local constant 10 (it shadows the input option). *warn is
 * incremented when the projected-vertex denominator falls below 1e-6 and the
 * actual vertices are used instead (the reference only warns). */
double ao_pot_source(const double *g, double dou, const double pos[3], int *warn) {
  int nv = (int)g[G_NVER];
  double rv[3] = {pos[0] - g[G_CEN], pos[1] - g[G_CEN + 1], pos[2] - g[G_CEN + 2]};
  double radius = ao_norm(rv);
  if (radius > 10.0 * maxedge(g)) return g[G_AREA] / radius;
  const double *e3 = &g[G_NOR];
  double zQ = ao_dot(rv, e3);
  double Qp[3] = {pos[0] - zQ * e3[0], pos[1] - zQ * e3[1], pos[2] - zQ * e3[2]};
  double sou = 0.0;
  for (int i1 = 0; i1 < nv; ++i1) {
    int ip1 = (nv == 3) ? next_tri[i1] : next_qua[i1];
    double d1[3], d2[3];
    for (int e = 0; e < 3; ++e) {
      d1[e] = pos[e] - g[G_VERP + 3 * i1 + e];
      d2[e] = pos[e] - g[G_VERP + 3 * ip1 + e];
    }
    double R1 = ao_norm(d1), R2 = ao_norm(d2);
    double el = g[G_ELEN + i1];
    double den = R1 + R2 - el;
    if (den < 1e-6) {
      if (warn) (*warn)++;
      for (int e = 0; e < 3; ++e) {
        d1[e] = pos[e] - g[G_VER + 3 * i1 + e];
        d2[e] = pos[e] - g[G_VER + 3 * ip1 + e];
      }
      R1 = ao_norm(d1);
      R2 = ao_norm(d2);
      den = R1 + R2 - el;
    }
    double souLog = log((R1 + R2 + el) / den);
    double q[3], cr[3];
    for (int e = 0; e < 3; ++e) q[e] = Qp[e] - g[G_VERP + 3 * i1 + e];
    ao_cross(q, &g[G_EVEC + 3 * i1], cr);
    double vi = -ao_dot(cr, e3) / el;
    sou += vi * souLog;
  }
  return sou - zQ * dou;
}

/* velocity_calc_doublet, mod_doublet.f90:178-248 (vortex ring with the
 * Rankine-like edge core and the cut-off). */
void ao_vel_doublet(const double *g, const double pos[3], double ff_ratio, double r_rankine,
                    double r_cutoff, double v[3]) {
  int nv = (int)g[G_NVER];
  double rv[3] = {pos[0] - g[G_CEN], pos[1] - g[G_CEN + 1], pos[2] - g[G_CEN + 2]};
  double radius = ao_norm(rv);
  v[0] = v[1] = v[2] = 0.0;
  if (radius > ff_ratio * maxedge(g)) {
    double r2 = radius * radius, r5 = r2 * r2 * radius;
    double rn = ao_dot(rv, &g[G_NOR]);
    double rati = 3.0 * g[G_AREA] * rn / r5;
    double phix = rati * ao_dot(rv, &g[G_TANG]);
    double phiy = rati * ao_dot(rv, &g[G_TANG + 3]);
    double pdou = g[G_AREA] * (-r2 + 3.0 * rn * rn) / r5;
    for (int e = 0; e < 3; ++e)
      v[e] = g[G_TANG + e] * phix + g[G_TANG + 3 + e] * phiy + g[G_NOR + e] * pdou;
    return;
  }
  for (int i1 = 0; i1 < nv; ++i1) {
    int ip1 = (nv == 3) ? next_tri[i1] : next_qua[i1];
    const double *eu = &g[G_EUNI + 3 * i1];
    double el = g[G_ELEN + i1];
    double av[3], bv[3];
    for (int e = 0; e < 3; ++e) {
      av[e] = pos[e] - g[G_VER + 3 * i1 + e];
      bv[e] = pos[e] - g[G_VER + 3 * ip1 + e];
    }
    double ai = ao_dot(av, eu);
    double R1 = ao_norm(av), R2 = ao_norm(bv);
    double hv[3];
    for (int e = 0; e < 3; ++e) hv[e] = av[e] - ai * eu[e];
    double hi = ao_norm(hv);
    double cr[3];
    ao_cross(eu, hv, cr);
    if (hi > el * r_rankine) {
      double f = ((el - ai) / R2 + ai / R1) / (hi * hi);
      for (int e = 0; e < 3; ++e) v[e] += f * cr[e];
    } else if (R1 > el * r_cutoff && R2 > el * r_cutoff) {
      double rr = r_rankine * el;
      double f = ((el - ai) / R2 + ai / R1) / (rr * rr);
      for (int e = 0; e < 3; ++e) v[e] += f * cr[e];
    }
  }
}

/* Static AIC of a surface-panel system (build_row_static_surfpan with
 * compute_pot_surfpan): A(i,j) = -dou_j(cen_i) (A_ii = 2 pi), B(i,j) = sou.
 * Column-major with leading dimensions lda, ldb. Returns the number of
 * degenerate-denominator warnings. */
int ao_assemble_surfpan(int n, const double *G, double ff_ratio_dou, double *A, int lda,
                        double *B, int ldb) {
  int warn = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : warn)
  for (int j = 0; j < n; ++j) {
    const double *gj = &G[(size_t)G_SIZE * j];
    for (int i = 0; i < n; ++i) {
      const double *pos = &G[(size_t)G_SIZE * i + G_CEN];
      double dou = (i != j) ? ao_pot_doublet(gj, pos, ff_ratio_dou) : -2.0 * AO_PI;
      A[(size_t)i + (size_t)lda * j] = -dou;
      int w = 0;
      double sou = ao_pot_source(gj, dou, pos, &w);
      warn += w;
      if (B) B[(size_t)i + (size_t)ldb * j] = sou;
    }
  }
  return warn;
}

/* Selected rows of the same matrices (the rows of build_row_static_surfpan
 * for the panels rows[0..nr)): Ar(r,j) = A(rows[r],j), Br(r,j) =
 * B(rows[r],j), both nr x n column-major with leading dimension nr. */
int ao_assemble_surfpan_rows(int n, const double *G, double ff_ratio_dou, int nr, const int *rows,
                             double *Ar, double *Br) {
  int warn = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : warn)
  for (int j = 0; j < n; ++j) {
    const double *gj = &G[(size_t)G_SIZE * j];
    for (int r = 0; r < nr; ++r) {
      const int i = rows[r];
      const double *pos = &G[(size_t)G_SIZE * i + G_CEN];
      double dou = (i != j) ? ao_pot_doublet(gj, pos, ff_ratio_dou) : -2.0 * AO_PI;
      Ar[(size_t)r + (size_t)nr * j] = -dou;
      int w = 0;
      double sou = ao_pot_source(gj, dou, pos, &w);
      warn += w;
      Br[(size_t)r + (size_t)nr * j] = sou;
    }
  }
  return warn;
}

/* Static AIC of a vortex-lattice system (build_row_static_vortlatt with
 * compute_psi_vortlatt): A(i,j) = v_ring_j(cen_i).nor_i, B(i,i) = 4 pi. */
void ao_assemble_vortlatt(int n, const double *G, double ff_ratio_dou, double r_rankine,
                          double r_cutoff, double *A, int lda, double *B, int ldb) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < n; ++j) {
    const double *gj = &G[(size_t)G_SIZE * j];
    for (int i = 0; i < n; ++i) {
      double v[3];
      ao_vel_doublet(gj, &G[(size_t)G_SIZE * i + G_CEN], ff_ratio_dou, r_rankine, r_cutoff, v);
      A[(size_t)i + (size_t)lda * j] = ao_dot(v, &G[(size_t)G_SIZE * i + G_NOR]);
      if (B) B[(size_t)i + (size_t)ldb * j] = (i == j) ? 4.0 * AO_PI : 0.0;
    }
  }
}

/* Static right-hand side, build_row_surfpan / build_row_vortlatt static part
 * (mod_surfpan.f90:400-404): b(i) = sum_j B(i,j) * nor_j.(-u_inf - uvort_j). */
void ao_rhs_static(int n, const double *G, const double *B, int ldb, const double uinf[3],
                   const double *uvort, double *b) {
  double *s = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  for (int j = 0; j < n; ++j) {
    double w[3];
    for (int e = 0; e < 3; ++e) w[e] = -uinf[e] - (uvort ? uvort[3 * j + e] : 0.0);
    s[j] = ao_dot(&G[(size_t)G_SIZE * j + G_NOR], w);
  }
#pragma omp parallel for
  for (int i = 0; i < n; ++i) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += B[(size_t)i + (size_t)ldb * j] * s[j];
    b[i] = acc;
  }
  free(s);
}

/* LU with partial pivoting, LAPACK dgetrf semantics (column-major, 1-based
 * ipiv, info > 0 for an exactly zero pivot; pivot = first maximum of |a|).
 * Right-looking, unblocked: O(n^3), for test sizes. */
int ao_dgetrf(int m, int n, double *A, int lda, int *ipiv) {
  int info = 0;
  int mn = m < n ? m : n;
  for (int j = 0; j < mn; ++j) {
    int p = j;
    double amax = fabs(A[j + (size_t)lda * j]);
    for (int i = j + 1; i < m; ++i) {
      double a = fabs(A[i + (size_t)lda * j]);
      if (a > amax) { amax = a; p = i; }
    }
    ipiv[j] = p + 1;
    if (A[p + (size_t)lda * j] != 0.0) {
      if (p != j)
        for (int k = 0; k < n; ++k) {
          double t = A[j + (size_t)lda * k];
          A[j + (size_t)lda * k] = A[p + (size_t)lda * k];
          A[p + (size_t)lda * k] = t;
        }
      double r = 1.0 / A[j + (size_t)lda * j];
      for (int i = j + 1; i < m; ++i) A[i + (size_t)lda * j] *= r;
    } else if (info == 0) {
      info = j + 1;
    }
#pragma omp parallel for schedule(static) if ((m - j) * (n - j) > 65536)
    for (int k = j + 1; k < n; ++k) {
      double u = A[j + (size_t)lda * k];
      if (u != 0.0)
        for (int i = j + 1; i < m; ++i) A[i + (size_t)lda * k] -= A[i + (size_t)lda * j] * u;
    }
  }
  return info;
}

/* dgetrs('N'): apply row interchanges, then L (unit) and U solves. */
void ao_dgetrs(int n, int nrhs, const double *A, int lda, const int *ipiv, double *B, int ldb) {
  for (int r = 0; r < nrhs; ++r) {
    double *b = &B[(size_t)ldb * r];
    for (int i = 0; i < n; ++i) {
      int p = ipiv[i] - 1;
      if (p != i) { double t = b[i]; b[i] = b[p]; b[p] = t; }
    }
    for (int j = 0; j < n; ++j) {
      double x = b[j];
      for (int i = j + 1; i < n; ++i) b[i] -= A[i + (size_t)lda * j] * x;
    }
    for (int j = n - 1; j >= 0; --j) {
      b[j] /= A[j + (size_t)lda * j];
      double x = b[j];
      for (int i = 0; i < j; ++i) b[i] -= A[i + (size_t)lda * j] * x;
    }
  }
}

int ao_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
