/*
 * elem_oracle.c -- CPU restatement of the direct element <-> particle terms of
 * a DUST time step: the body / wake element velocity (and optional stretching)
 * at the vortex particles, and the particle / end-vortex velocity at the panel
 * centres (uvort).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product library never does.
 *
 * Reference files (paths relative to the DUST source tree):
 *   src/fmm/mod_octree.f90        apply_multipole element loops 1350-1401
 *   src/aeroel/mod_surfpan.f90    velocity_calc_sou_surfpan 268-357,
 *                                 gradient_calc_sou_surfpan 362-370 (zero),
 *                                 compute_vel_surfpan 641-658,
 *                                 compute_grad_surfpan 670-687,
 *                                 get_vort_vel_surfpan 962-977
 *   src/aeroel/mod_doublet.f90    velocity_calc_doublet 178-248,
 *                                 gradient_calc_doublet 254-305
 *   src/aeroel/mod_vortlatt.f90   compute_vel/grad_vortlatt 403-445,
 *                                 get_vort_vel_vortlatt 676-693
 *   src/aeroel/mod_vortline.f90   compute_vel_vortline 126-170,
 *                                 compute_grad_vortline 174-236,
 *                                 calc_geo_data_vortline 238-258
 *   src/aeroel/mod_vortpart.f90   compute_vel_vortpart 139-153, kernel_coeffs
 *                                 260-271 (Rosenhead)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "aic_oracle.h"

/* velocity_calc_sou_surfpan (mod_surfpan.f90:268-357): the routine's own
 * far-field ratio is the local parameter 10; its near field calls
 * potential_calc_doublet, whose ratio is FarFieldRatioDoublet. */
void ao_vel_source(const double *g, const double pos[3], double ff_ratio_dou, double v[3]) {
  int nv = (int)g[G_NVER];
  const double *t1 = &g[G_TANG], *t2 = &g[G_TANG + 3], *nor = &g[G_NOR];
  double rv[3] = {pos[0] - g[G_CEN], pos[1] - g[G_CEN + 1], pos[2] - g[G_CEN + 2]};
  double radius = ao_norm(rv);
  double phix, phiy, pdou;
  if (radius > 10.0 * maxedge(g)) {
    double r3 = pow(radius, 3.0);
    phix = -g[G_AREA] * ao_dot(rv, t1) / r3;
    phiy = -g[G_AREA] * ao_dot(rv, t2) / r3;
    pdou = -g[G_AREA] * ao_dot(rv, nor) / r3;
  } else {
    phix = 0.0;
    phiy = 0.0;
    for (int i1 = 0; i1 < nv; ++i1) {
      int ip1 = (nv == 3) ? next_tri[i1] : next_qua[i1];
      double a[3], b[3];
      for (int e = 0; e < 3; ++e) {
        a[e] = pos[e] - g[G_VERP + 3 * i1 + e];
        b[e] = pos[e] - g[G_VERP + 3 * ip1 + e];
      }
      double R1 = ao_norm(a), R2 = ao_norm(b), el = g[G_ELEN + i1];
      double souLog = (R1 + R2 - el < 1e-12) ? 0.0 : log((R1 + R2 + el) / (R1 + R2 - el));
      phix += g[G_SINT + i1] * souLog;
      phiy -= g[G_COST + i1] * souLog;
    }
    pdou = ao_pot_doublet(g, pos, ff_ratio_dou);
    phix = -phix;
    phiy = -phiy;
    pdou = -pdou;
  }
  for (int e = 0; e < 3; ++e) v[e] = t1[e] * phix + t2[e] * phiy + nor[e] * pdou;
}

/* One edge of gradient_calc_doublet / compute_grad_vortline (the two share
 * the expression): the edge from x1 to x2 with unit vector l, core `del`.
 * grad(3,3) is column-major, grad[r + 3 c], and is accumulated. */
static void edge_grad(const double *l, const double *x1, const double *x2, const double pos[3],
                      double del, double *grad) {
  double R1[3], R2[3], a1[3], a2[3];
  for (int e = 0; e < 3; ++e) {
    R1[e] = pos[e] - x1[e];
    R2[e] = pos[e] - x2[e];
  }
  ao_cross(l, R1, a1);
  ao_cross(l, R2, a2);
  double a = ao_norm(a1);
  double a2del2 = pow(a, 2.0) + pow(del, 2.0);
  /* lx(:,1) = (0, l3, -l2), lx(:,2) = (-l3, 0, l1), lx(:,3) = (l2, -l1, 0) */
  double lx[9] = {0.0, l[2], -l[1], -l[2], 0.0, l[0], l[1], -l[0], 0.0};
  for (int t = 0; t < 2; ++t) {
    const double *R = t ? R2 : R1, *av = t ? a2 : a1;
    double nR = ao_norm(R), lR = ao_dot(l, R);
    double fac = 1.0 / (pow(a2del2, 1.5) * nR);
    double coef = 1.0 / a - 3.0 * a / a2del2;
    double sgn = t ? -1.0 : 1.0;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) {
        double aalx = 0.0; /* matmul(aa, lx)(r,c) */
        for (int k = 0; k < 3; ++k) aalx += av[r] * av[k] * lx[k + 3 * c];
        double m = (a * lx[r + 3 * c] + coef * aalx) * lR +
                   a * (av[r] * l[c] - av[r] * R[c] * lR / pow(nR, 2.0));
        grad[r + 3 * c] += sgn * fac * m;
      }
  }
}

/* gradient_calc_doublet (mod_doublet.f90:254-305); del = VortexRad. */
void ao_grad_doublet(const double *g, const double pos[3], double del, double *grad) {
  int nv = (int)g[G_NVER];
  memset(grad, 0, sizeof(double) * 9);
  for (int i = 0; i < nv; ++i) {
    int ip1 = (nv == 3) ? next_tri[i] : next_qua[i];
    edge_grad(&g[G_EUNI + 3 * i], &g[G_VER + 3 * i], &g[G_VER + 3 * ip1], pos, del, grad);
  }
}

/* compute_vel_vortline without the strength (mod_vortline.f90:126-170);
 * ver(3,2) is the line. */
void ao_vel_vortline(const double *ver, const double pos[3], double r_rankine, double r_cutoff,
                     double v[3]) {
  double ev[3], eu[3];
  for (int e = 0; e < 3; ++e) ev[e] = ver[3 + e] - ver[e];
  double el = ao_norm(ev);
  for (int e = 0; e < 3; ++e) eu[e] = ev[e] / el;
  double av[3], bv[3], hv[3], cr[3];
  for (int e = 0; e < 3; ++e) {
    av[e] = pos[e] - ver[e];
    bv[e] = pos[e] - ver[3 + e];
  }
  double ai = ao_dot(av, eu);
  double R1 = ao_norm(av), R2 = ao_norm(bv);
  for (int e = 0; e < 3; ++e) hv[e] = av[e] - ai * eu[e];
  double hi = ao_norm(hv);
  ao_cross(eu, hv, cr);
  v[0] = v[1] = v[2] = 0.0;
  if (hi > el * r_rankine) {
    double f = ((el - ai) / R2 + ai / R1) / pow(hi, 2.0);
    for (int e = 0; e < 3; ++e) v[e] = f * cr[e];
  } else if (R1 > el * r_cutoff && R2 > el * r_cutoff) {
    double rr = r_rankine * el;
    double f = ((el - ai) / R2 + ai / R1) / pow(rr, 2.0);
    for (int e = 0; e < 3; ++e) v[e] = f * cr[e];
  }
}

/* compute_grad_vortline without the strength (mod_vortline.f90:174-236). */
void ao_grad_vortline(const double *ver, const double pos[3], double del, double *grad) {
  double ev[3], eu[3];
  for (int e = 0; e < 3; ++e) ev[e] = ver[3 + e] - ver[e];
  double el = ao_norm(ev);
  for (int e = 0; e < 3; ++e) eu[e] = ev[e] / el;
  memset(grad, 0, sizeof(double) * 9);
  edge_grad(eu, ver, ver + 3, pos, del, grad);
}

/* Element kinds of one group (the element lists of apply_multipole). */
enum { AO_SURFPAN = 0, AO_RING = 1, AO_LINE = 2 };

/* One element list of apply_multipole (mod_octree.f90:1350-1401) at npts
 * points: vel += v / (4 pi) element by element; when `stretch` is given,
 * stretch += transpose(grad) alpha / (4 pi).  Polygons come as panel records
 * G (ao_geo), lines as ver(3,2,n).  mag NULL: no doublet / line term and no
 * gradient (a line whose strength is not associated induces nothing);
 * sigma (surface panels; NULL: no source term) is the source strength
 * nor.(ub - wind - uvort) of compute_vel_surfpan. */
void ao_elem_group(int kind, int n, const double *G, const double *lver, const double *mag,
                   const double *sigma, double ff_ratio_dou, double r_rankine, double r_cutoff,
                   double vortex_rad, int npts, const double *pos, const double *alpha,
                   double *vel, double *stretch) {
  const double f4pi = 4.0 * AO_PI;
#pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < npts; ++p) {
    const double *x = &pos[3 * p];
    for (int j = 0; j < n; ++j) {
      double v[3] = {0.0, 0.0, 0.0};
      double m = mag ? mag[j] : 0.0;
      if (kind == AO_LINE) {
        if (mag) {
          double vd[3];
          ao_vel_vortline(&lver[6 * j], x, r_rankine, r_cutoff, vd);
          for (int e = 0; e < 3; ++e) v[e] = vd[e] * m;
        }
      } else {
        const double *g = &G[(size_t)G_SIZE * j];
        double vd[3] = {0.0, 0.0, 0.0}, vs[3] = {0.0, 0.0, 0.0};
        if (mag) ao_vel_doublet(g, x, ff_ratio_dou, r_rankine, r_cutoff, vd);
        if (kind == AO_SURFPAN && sigma) ao_vel_source(g, x, ff_ratio_dou, vs);
        double s = (kind == AO_SURFPAN && sigma) ? sigma[j] : 0.0;
        for (int e = 0; e < 3; ++e) v[e] = vd[e] * m - vs[e] * s;
      }
      for (int e = 0; e < 3; ++e) vel[3 * p + e] = vel[3 * p + e] + v[e] / f4pi;
    }
    if (stretch && mag) {
      for (int j = 0; j < n; ++j) {
        double gr[9];
        if (kind == AO_LINE) {
          ao_grad_vortline(&lver[6 * j], x, vortex_rad, gr);
        } else {
          ao_grad_doublet(&G[(size_t)G_SIZE * j], x, vortex_rad, gr);
        }
        double m = mag[j];
        const double *al = &alpha[3 * p];
        for (int c = 0; c < 3; ++c) {
          double s = 0.0; /* matmul(transpose(grad), alpha)(c) */
          for (int r = 0; r < 3; ++r) s += gr[r + 3 * c] * m * al[r];
          stretch[3 * p + c] += s / f4pi;
        }
      }
    }
  }
}

/* get_vort_vel_surfpan / _vortlatt (mod_surfpan.f90:962-977,
 * mod_vortlatt.f90:676-693): uvort(:,i) at the m centres, from the np vortex
 * particles (compute_vel_vortpart, Rosenhead core r_vortex of the source)
 * followed by the nl end-vortex lines (compute_vel_vortline). */
void ao_vort_vel(int m, const double *cen, int np, const double *pcen, const double *pdir,
                 const double *pmag, const double *prv, int nl, const double *lver,
                 const double *lmag, double r_rankine, double r_cutoff, double *uvort) {
  const double f4pi = 4.0 * AO_PI;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m; ++i) {
    const double *x = &cen[3 * i];
    double u[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < np; ++j) {
      double dist[3], cr[3];
      for (int e = 0; e < 3; ++e) dist[e] = x[e] - pcen[3 * j + e];
      double distn = ao_norm(dist);
      double R = sqrt(pow(distn, 2.0) + pow(prv[j], 2.0));
      double c = -1.0 / pow(R, 3.0);
      ao_cross(dist, &pdir[3 * j], cr);
      for (int e = 0; e < 3; ++e) u[e] = u[e] + c * cr[e] * pmag[j] / f4pi;
    }
    for (int j = 0; j < nl; ++j) {
      if (!lmag) break;
      double vd[3];
      ao_vel_vortline(&lver[6 * j], x, r_rankine, r_cutoff, vd);
      for (int e = 0; e < 3; ++e) u[e] = u[e] + vd[e] * lmag[j] / f4pi;
    }
    for (int e = 0; e < 3; ++e) uvort[3 * i + e] = u[e];
  }
}
