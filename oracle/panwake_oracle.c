/*
 * panwake_oracle.c -- CPU restatement of DUST's panel wake: the trailing-edge
 * rows of initialize_wake / complete_wake, the panel part of update_wake,
 * join_first_panels / get_joined_intensity and the particles shed from the
 * last panel row (classical and refined).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product library never does.
 *
 * Reference file: src/aeroel/mod_wake.f90
 *   initialize_wake 512-546, update_wake 884-955 and 1133-1187,
 *   complete_wake 1250-1288 and 1798-1969 and 2089-2103,
 *   join_first_panels 2199-2313, get_joined_intensity 2315-2345.
 *
 * Storage (0-based): points and velocities [row][point][3] with rows
 * 0..nmax (the reference's 1..nmax_pan+1), intensities [row][stripe], the
 * junction history [row][end][side].  The tables (start points, neighbours,
 * generating elements) are 1-based as the reference's.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int ns, np, nmax, len, full, nend;
  const int *start, *neigh, *neigh_o, *gen; /* (2, ns) each */
  const int *ends;                          /* (nend) 0-based stripes */
  int *joined;                              /* (nmax+1, nend, 2) 1-based stripe or 0 */
  double *pts, *pvel, *wstart, *pend, *idou, *end_idou, *last_idou;
  double dt;
  int ndt;
  double min_vel;
  int join_te;
  double join_fact;
  int refine, k_refine;
  double kvr, vr;
} ao_pw;

static double n3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static double *P(const ao_pw *w, int row, int ip) { return &w->pts[3 * ((size_t)row * w->np + ip)]; }
static double *V(const ao_pw *w, int row, int ip) { return &w->pvel[3 * ((size_t)row * w->np + ip)]; }
static int sp(const ao_pw *w, int iw, int s) { return w->start[2 * iw + s] - 1; }

/* join_first_panels (2199-2313) */
static void join_first_panels(ao_pw *w) {
  for (int i = 0; i < w->nend; ++i) {
    int iw = w->ends[i];
    const double *sp1 = &w->wstart[3 * sp(w, iw, 0)], *sp2 = &w->wstart[3 * sp(w, iw, 1)];
    double t[3];
    for (int e = 0; e < 3; ++e) t[e] = sp1[e] - sp2[e];
    double l1 = n3(t);
    for (int j = i + 1; j < w->nend; ++j) {
      int ir = w->ends[j];
      const double *q1 = &w->wstart[3 * sp(w, ir, 0)], *q2 = &w->wstart[3 * sp(w, ir, 1)];
      for (int e = 0; e < 3; ++e) t[e] = q1[e] - q2[e];
      double l2 = n3(t);
      double tol = w->join_fact * (l1 < l2 ? l1 : l2);
      double d;
      int icase = 1;
      for (int e = 0; e < 3; ++e) t[e] = q1[e] - sp1[e];
      double lmin = n3(t);
      for (int e = 0; e < 3; ++e) t[e] = q2[e] - sp1[e];
      if ((d = n3(t)) < lmin) { lmin = d; icase = 2; }
      for (int e = 0; e < 3; ++e) t[e] = q1[e] - sp2[e];
      if ((d = n3(t)) < lmin) { lmin = d; icase = 3; }
      for (int e = 0; e < 3; ++e) t[e] = q2[e] - sp2[e];
      if ((d = n3(t)) < lmin) { lmin = d; icase = 4; }
      const double *a = (icase <= 2) ? sp1 : sp2;
      const double *b = (icase == 1 || icase == 3) ? q1 : q2;
      int si = (icase <= 2) ? 0 : 1, sj = (icase == 1 || icase == 3) ? 0 : 1;
      for (int e = 0; e < 3; ++e) t[e] = a[e] - b[e];
      if (n3(t) < tol) {
        int pi = sp(w, iw, si), pj = sp(w, ir, sj);
        double m[3];
        for (int e = 0; e < 3; ++e) m[e] = (a[e] + b[e]) / 2.0;
        memcpy(P(w, 0, pi), m, sizeof m);
        memcpy(P(w, 0, pj), m, sizeof m);
        for (int e = 0; e < 3; ++e) m[e] = (P(w, 1, pi)[e] + P(w, 1, pj)[e]) / 2.0;
        memcpy(P(w, 1, pi), m, sizeof m);
        memcpy(P(w, 1, pj), m, sizeof m);
        w->joined[2 * i + si] = ir + 1;
        w->joined[2 * j + sj] = iw + 1;
      }
    }
  }
}

/* Rows 1 and 2 (initialize_wake 512-546 with first = 1; complete_wake
 * 1250-1288 with first = 0, which also moves the junction history back). */
void ao_pw_start(ao_pw *w, int first, const double *spts, const double *gdir, const double *scal,
                 const double *vte, const double uinf[3]) {
  double dtw = w->dt * w->ndt;
  for (int ip = 0; ip < w->np; ++ip) {
    memcpy(&w->wstart[3 * ip], &spts[3 * ip], 3 * sizeof(double));
    memcpy(P(w, 0, ip), &spts[3 * ip], 3 * sizeof(double));
    double d[3], r[3];
    for (int e = 0; e < 3; ++e) {
      d[e] = gdir[3 * ip + e];
      r[e] = uinf[e] - vte[3 * ip + e];
    }
    double wn = n3(r);
    for (int e = 0; e < 3; ++e) {
      double x;
      if (wn > w->min_vel)
        x = spts[3 * ip + e] + d[e] * scal[ip] * wn * dtw / n3(d);
      else if (first)
        x = spts[3 * ip + e] + d[e] * scal[ip] * w->min_vel * dtw;
      else
        x = spts[3 * ip + e] + d[e] * scal[ip] * w->min_vel * dtw / n3(d);
      P(w, 1, ip)[e] = x;
    }
  }
  if (w->join_te) {
    if (!first) {
      size_t blk = 2 * (size_t)w->nend;
      for (int r = w->nmax; r >= 1; --r)
        memcpy(&w->joined[r * blk], &w->joined[(r - 1) * blk], blk * sizeof(int));
      memset(w->joined, 0, blk * sizeof(int));
    }
    join_first_panels(w);
  }
}

/* update_wake, panel part (884-955, 1133-1187); vel(3, len*np) at the rows
 * 2..len+1 of the points. */
void ao_pw_update(ao_pw *w, const double *mu, const double *vel) {
  if (w->len == w->nmax) w->full = 1;
  for (int iw = 0; iw < w->ns; ++iw) {
    int g1 = w->gen[2 * iw], g2 = w->gen[2 * iw + 1];
    w->idou[iw] = g2 > 0 ? mu[g1 - 1] - mu[g2 - 1] : mu[g1 - 1];
  }
  double *old = malloc(sizeof(double) * 3 * (size_t)(w->nmax + 1) * w->np);
  memcpy(old, w->pts, sizeof(double) * 3 * (size_t)(w->nmax + 1) * w->np);
  int np1 = w->len + 1 + (w->full ? 0 : 1);
  double dtw = w->dt * w->ndt;
  for (int row = 3; row <= np1; ++row)
    for (int ip = 0; ip < w->np; ++ip) {
      const double *xo = &old[3 * ((size_t)(row - 2) * w->np + ip)];
      const double *v = &vel[3 * ((size_t)(row - 3) * w->np + ip)];
      for (int e = 0; e < 3; ++e) {
        V(w, row - 1, ip)[e] = v[e];
        P(w, row - 1, ip)[e] = xo[e] + v[e] * (row != 3 ? dtw : w->dt);
      }
    }
  if (w->full)
    for (int ip = 0; ip < w->np; ++ip) {
      const double *xo = &old[3 * ((size_t)w->len * w->np + ip)];
      const double *v = &vel[3 * ((size_t)(w->len - 1) * w->np + ip)];
      for (int e = 0; e < 3; ++e) w->pend[3 * ip + e] = xo[e] + v[e] * dtw;
    }
  free(old);
  if (!w->full) w->len += 1;
  for (int iw = 0; iw < w->ns; ++iw) w->end_idou[iw] = w->idou[(size_t)(w->len - 1) * w->ns + iw];
  for (int row = w->len; row >= 2; --row)
    for (int iw = 0; iw < w->ns; ++iw)
      w->idou[(size_t)(row - 1) * w->ns + iw] = w->idou[(size_t)(row - 2) * w->ns + iw];
}

/* normal direction (unnormalised) of the first-row panel of stripe iw */
static void first_nor(const ao_pw *w, int iw, double n[3]) {
  const double *v1 = P(w, 0, sp(w, iw, 0)), *v2 = P(w, 0, sp(w, iw, 1)),
               *v3 = P(w, 1, sp(w, iw, 1)), *v4 = P(w, 1, sp(w, iw, 0));
  double a[3], b[3];
  for (int e = 0; e < 3; ++e) {
    a[e] = v3[e] - v1[e];
    b[e] = v4[e] - v2[e];
  }
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
}

/* get_joined_intensity (2315-2345) */
static double joined_intensity(const ao_pw *w, int iw, int side) {
  int iend;
  for (iend = 0; iend < w->nend; ++iend)
    if (w->ends[iend] == iw) break;
  int ineigh = w->joined[(size_t)w->nmax * 2 * w->nend + 2 * iend + side] - 1;
  double mi = w->idou[(size_t)(w->len - 1) * w->ns + iw];
  if (ineigh < 0) return mi;
  double a[3], b[3];
  first_nor(w, iw, a);
  first_nor(w, ineigh, b);
  double orient = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) >= 0.0 ? 1.0 : -1.0;
  return (mi - orient * w->idou[(size_t)(w->len - 1) * w->ns + ineigh]) / 2.0;
}

/* Shedding when the panel wake is full (1798-1969): new particles go, in
 * order, into the lowest free slots of the pool (free[i] = 1) if inside the
 * box.  Returns 0, or -4 when the free slots run out (the reference stops). */
int ao_pw_shed(ao_pw *w, int nmax, unsigned char *freef, double *cen, double *dir, double *mag,
               double *rv, double *vel, const double bmin[3], const double bmax[3], int *n_prt) {
  if (!w->full) return 0;
  int k = 0;
  const int R = w->nmax;
  for (int iw = 0; iw < w->ns; ++iw) {
    int p1 = sp(w, iw, 0), p2 = sp(w, iw, 1);
    const double *P1 = P(w, R, p1), *P2 = P(w, R, p2), *E1 = &w->pend[3 * p1],
                 *E2 = &w->pend[3 * p2];
    double partvec[3] = {0.0, 0.0, 0.0}, d[3], ave;
    for (int e = 0; e < 3; ++e) d[e] = P1[e] - E1[e];
    double chord = n3(d);
    if (w->neigh[2 * iw] > 0)
      ave = (w->end_idou[iw] - (double)w->neigh_o[2 * iw] * w->end_idou[w->neigh[2 * iw] - 1]) / 2.0;
    else
      ave = w->join_te ? joined_intensity(w, iw, 0) : w->end_idou[iw];
    for (int e = 0; e < 3; ++e) partvec[e] += d[e] * ave;
    for (int e = 0; e < 3; ++e) d[e] = -P2[e] + E2[e];
    if (w->neigh[2 * iw + 1] > 0)
      ave = (w->end_idou[iw] -
             (double)w->neigh_o[2 * iw + 1] * w->end_idou[w->neigh[2 * iw + 1] - 1]) / 2.0;
    else
      ave = w->join_te ? joined_intensity(w, iw, 1) : w->end_idou[iw];
    for (int e = 0; e < 3; ++e) partvec[e] += d[e] * ave;
    for (int e = 0; e < 3; ++e) d[e] = E1[e] - E2[e];
    double span = n3(d);
    ave = w->end_idou[iw] - w->last_idou[iw];
    w->last_idou[iw] = w->end_idou[iw];
    for (int e = 0; e < 3; ++e) partvec[e] += d[e] * ave;
    const double *W1 = V(w, R, p1), *W2 = V(w, R, p2);
    int nch = 1, nsp = 1;
    double stc = 0.0, sts = 0.0;
    if (w->refine) {
      if (chord >= span) {
        nch = (int)ceil(chord / (span / (double)w->k_refine));
        nsp = w->k_refine;
        stc = chord / nch;
        sts = span / nsp;
      } else {
        nch = w->k_refine;
        nsp = (int)ceil(span / (chord / (double)w->k_refine));
        stc = chord / nch;
        sts = span / nsp;
      }
    }
    double pvn = n3(partvec);
    for (int ic = 1; ic <= nch; ++ic)
      for (int is = 1; is <= nsp; ++is) {
        double pos[3], area, m;
        if (w->refine) {
          for (int e = 0; e < 3; ++e)
            pos[e] = P1[e] + ((double)ic - 0.5) / (double)nch * (E1[e] - P1[e]) +
                     ((double)is - 0.5) / (double)nsp * (P2[e] - P1[e]);
          area = stc * sts;
          m = pvn / (double)(nsp * nch);
        } else {
          double a[3], b[3], cr[3];
          for (int e = 0; e < 3; ++e) {
            pos[e] = (E1[e] + E2[e] + P1[e] + P2[e]) / 4.0;
            a[e] = E1[e] - P2[e];
            b[e] = E2[e] - P2[e];
          }
          cr[0] = a[1] * b[2] - a[2] * b[1];
          cr[1] = a[2] * b[0] - a[0] * b[2];
          cr[2] = a[0] * b[1] - a[1] * b[0];
          area = n3(cr);
          m = pvn;
        }
        int in = 1;
        for (int e = 0; e < 3; ++e) in = in && pos[e] >= bmin[e] && pos[e] <= bmax[e];
        if (!in) continue;
        while (k < nmax && !freef[k]) ++k;
        if (k >= nmax) return -4;
        freef[k] = 0;
        *n_prt += 1;
        mag[k] = m;
        double thr = w->refine ? pvn : m;
        for (int e = 0; e < 3; ++e) {
          dir[3 * k + e] = thr > 1.0e-13 ? partvec[e] / pvn : partvec[e];
          cen[3 * k + e] = pos[e];
          vel[3 * k + e] = 0.5 * (W1[e] + W2[e]);
        }
        rv[k] = w->kvr >= 1e-10 ? w->kvr * sqrt(2.0 * area) : w->vr;
        ++k;
      }
  }
  return 0;
}

/* End vortices (2089-2103): the last row of points and their velocities;
 * the intensity is last_idou. */
void ao_pw_end_vortices(const ao_pw *w, double *lver, double *lvel) {
  int row = w->len;
  for (int iw = 0; iw < w->ns; ++iw)
    for (int s = 0; s < 2; ++s) {
      memcpy(&lver[6 * iw + 3 * s], P(w, row, sp(w, iw, s)), 3 * sizeof(double));
      memcpy(&lvel[6 * iw + 3 * s], V(w, row, sp(w, iw, s)), 3 * sizeof(double));
    }
}
