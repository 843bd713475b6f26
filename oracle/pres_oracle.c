/*
 * pres_oracle.c -- CPU restatement of the rotational part of the right-hand
 * side of DUST's pressure (Bernoulli) integral equation.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product library never does.
 *
 * Reference file: src/linsys/mod_pressure_equation.f90, assemble_pressure_sys
 *   (c.1) particles            171-187
 *   (c.2) end-vortex lines     191-203
 *   (c.3) wake panels          204-225
 * and press_normvel_der (396-462), the panels' bernoulli_source.
 * Each panel's b_pres receives the terms one at a time in the reference's
 * loop order (all particles first, in a loop of its own, then the lines and
 * the wake panels edge by edge).
 */
#include <math.h>

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* sum(e * (cross(d1, v1) / norm2(d1)**3 + cross(d2, v2) / norm2(d2)**3)) */
static double edge_sum(const double *p1, const double *p2, const double *e, const double *v1,
                       const double *v2, const double *x) {
  double d1[3], d2[3], c1[3], c2[3];
  for (int k = 0; k < 3; ++k) {
    d1[k] = p1[k] - x[k];
    d2[k] = p2[k] - x[k];
  }
  cross3(d1, v1, c1);
  cross3(d2, v2, c2);
  double n1 = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
  double n2 = sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
  double q1 = pow(n1, 3.0), q2 = pow(n2, 3.0);
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s += e[k] * (c1[k] / q1 + c2[k] / q2);
  return s;
}

/* b(m) += rotational terms at the centres cen(3,m).  Arrays as in
 * dust_pres_rhs_rot: particles pcen/pdir/pvel (3,np), pmag, prv (np); lines
 * lver/lvel (3,2,nl), lmag (nl, NULL = not associated); wake panels
 * wver/wvel (3,4,nw), wnver, wmag (nw). */
void ao_pres_rhs_rot(int m, const double *cen, int np, const double *pcen, const double *pdir,
                     const double *pmag, const double *prv, const double *pvel, int nl,
                     const double *lver, const double *lmag, const double *lvel, int nw,
                     const int *wnver, const double *wver, const double *wmag,
                     const double *wvel, double *b) {
#pragma omp parallel for schedule(dynamic, 32)
  for (int ie = 0; ie < m; ++ie) {
    const double *x = &cen[3 * ie];
    for (int iw = 0; iw < np; ++iw) {
      double d[3], c[3];
      for (int k = 0; k < 3; ++k) d[k] = pcen[3 * iw + k] - x[k];
      cross3(&pdir[3 * iw], &pvel[3 * iw], c);
      double s = d[0] * c[0] + d[1] * c[1] + d[2] * c[2];
      double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + prv[iw] * prv[iw]);
      b[ie] = b[ie] + s * pmag[iw] / pow(r, 3.0);
    }
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int ie = 0; ie < m; ++ie) {
    const double *x = &cen[3 * ie];
    if (lmag)
      for (int iw = 0; iw < nl; ++iw) {
        const double *v = &lver[6 * iw], *u = &lvel[6 * iw];
        double e[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        b[ie] = b[ie] - 0.5 * lmag[iw] * edge_sum(v, v + 3, e, u, u + 3, x);
      }
    for (int w = 0; w < nw; ++w) {
      int nv = wnver[w];
      const double *v = &wver[12 * w], *u = &wvel[12 * w];
      for (int is = 0; is < nv; ++is) {
        int nx = (is + 1) % nv;
        double e[3];
        for (int k = 0; k < 3; ++k) e[k] = v[3 * nx + k] - v[3 * is + k];
        b[ie] = b[ie] - 0.5 * wmag[w] * edge_sum(v + 3 * is, v + 3 * nx, e, u + 3 * is,
                                                 u + 3 * nx, x);
      }
    }
  }
}

/* press_normvel_der, mod_pressure_equation.f90:396-462: bernoulli_source of
 * every surface panel (arrays as in dust_press_normvel_der: nor, ub, sv,
 * svo (3,n); st (3,4,n); nv (n); ng (4,n) 0-based neighbour or -1 (none:
 * -2 U) or -2 (not a surface panel: nothing); R (3,3,n) column-major). */
void ao_press_normvel_der(int n, const double *nor, const double *ub, const double *sv,
                          const double *svo, const double *st, const int *nv, const int *ng,
                          const double *R, double dt, double *out) {
  for (int i = 0; i < n; ++i) {
    const double *ni = nor + 3 * i, *u = sv + 3 * i, *b = ub + 3 * i, *Ri = R + 9 * i;
    double dun = 0.0;
    for (int k = 0; k < 3; ++k) dun += ni[k] * (b[k] - svo[3 * i + k]);
    dun = dun / dt;
    double g[3] = {0.0, 0.0, 0.0};
    for (int e = 0; e < nv[i]; ++e) {
      int j = ng[4 * i + e];
      if (j == -2) continue;
      double d[3], dn = 0.0;
      for (int k = 0; k < 3; ++k) d[k] = j >= 0 ? sv[3 * j + k] - u[k] : -2.0 * u[k];
      for (int k = 0; k < 3; ++k) dn += ni[k] * d[k];
      const double *s = st + 12 * i + 3 * e;
      double v[3] = {s[0] * dn, s[1] * dn, s[2] * dn};
      for (int r = 0; r < 3; ++r) {
        double m = 0.0;
        for (int k = 0; k < 3; ++k) m += Ri[r + 3 * k] * v[k];
        g[r] += m;
      }
    }
    double div = 0.0;
    for (int e = 0; e < nv[i]; ++e) {
      int j = ng[4 * i + e];
      if (j == -2) continue;
      const double *s = st + 12 * i + 3 * e;
      double q = 0.0;
      for (int r = 0; r < 3; ++r) {
        double rs = 0.0;
        for (int k = 0; k < 3; ++k) rs += Ri[r + 3 * k] * s[k];
        q += rs * (j >= 0 ? sv[3 * j + r] - u[r] : -2.0 * u[r]);
      }
      div += q;
    }
    double gu = 0.0, un = 0.0;
    for (int k = 0; k < 3; ++k) gu += g[k] * b[k];
    for (int k = 0; k < 3; ++k) un += b[k] * ni[k];
    out[i] = dun - gu + div * un;
  }
}
