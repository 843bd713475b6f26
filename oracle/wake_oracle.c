/*
 * wake_oracle.c -- CPU restatement of the particle-pool steps of DUST's wake:
 * the particle update of complete_wake, the live list of prepare_wake and the
 * insertion of shed particles.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product library never does.
 *
 * Reference file: src/aeroel/mod_wake.f90
 *   complete_wake particle loop          1305-1355
 *   complete_wake shedding (classical)   1888-1967, ring branch 1974-2020
 *   prepare_wake pointer rebuild         826-835
 */
#include <math.h>

/* The particle loop of complete_wake over the pool (free slots skipped).
 * dt = sim_param%dt * ndt_update_wake, filt_eta = sim_param%filt_eta /
 * ndt_update_wake.  Returns the number of slots freed (n_prt decrements). */
int ao_wake_update(int nmax, unsigned char *freef, double *cen, double *dir, const double *mag,
                   const double *vel, double *stretch, const double *rotu,
                   const double *box_min, const double *box_max, double dt, double filt_eta,
                   int use_vs, int use_vd, int use_divfilt) {
  int freed = 0;
  for (int ip = 0; ip < nmax; ++ip) {
    if (freef[ip]) continue;
    double pos[3];
    for (int e = 0; e < 3; ++e) pos[e] = cen[3 * ip + e] + vel[3 * ip + e] * dt;
    int inside = 1;
    for (int e = 0; e < 3; ++e) inside &= (pos[e] >= box_min[e]);
    for (int e = 0; e < 3; ++e) inside &= (pos[e] <= box_max[e]);
    if (inside) {
      for (int e = 0; e < 3; ++e) cen[3 * ip + e] = pos[e];
      if (use_vs || use_vd) {
        double *s = &stretch[3 * ip], *d = &dir[3 * ip];
        double m = mag[ip];
        if (use_divfilt) {
          const double *r = &rotu[3 * ip];
          double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
          for (int e = 0; e < 3; ++e) s[e] = s[e] - filt_eta * (d[e] * m - r[e] * m / rn);
        }
        double a[3];
        for (int e = 0; e < 3; ++e) a[e] = d[e] * m + s[e] * dt;
        double an = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        if (an != 0.0)
          for (int e = 0; e < 3; ++e) d[e] = a[e] / an;
      }
    } else {
      freef[ip] = 1;
      ++freed;
    }
  }
  return freed;
}

/* prepare_wake: the non-free slots in slot order (0-based); returns the count. */
int ao_wake_live(int nmax, const unsigned char *freef, int *live) {
  int n = 0;
  for (int ir = 0; ir < nmax; ++ir)
    if (!freef[ir]) live[n++] = ir;
  return n;
}

/* Shedding insertion: new particles in order; one inside the box takes the
 * first free slot at or after the cursor k (which starts at the first slot
 * and moves past every slot taken).  Returns the number inserted, or -1 when
 * a particle finds no free slot (the reference stops there). */
int ao_wake_shed(int nmax, unsigned char *freef, double *cen, double *dir, double *mag, double *rv,
                 double *vel, int knew, const double *ncen, const double *npv, const double *nrv,
                 const double *nvel, const double *box_min, const double *box_max) {
  int k = 0, added = 0;
  for (int j = 0; j < knew; ++j) {
    const double *pos = &ncen[3 * j];
    int inside = 1;
    for (int e = 0; e < 3; ++e) inside &= (pos[e] >= box_min[e]) & (pos[e] <= box_max[e]);
    if (!inside) continue;
    int ip;
    for (ip = k; ip < nmax; ++ip) {
      if (freef[ip]) {
        freef[ip] = 0;
        k = ip + 1;
        ++added;
        const double *pv = &npv[3 * j];
        double m = sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]);
        mag[ip] = m;
        for (int e = 0; e < 3; ++e) dir[3 * ip + e] = (m > 1.0e-13) ? pv[e] / m : pv[e];
        for (int e = 0; e < 3; ++e) cen[3 * ip + e] = pos[e];
        rv[ip] = nrv[j];
        for (int e = 0; e < 3; ++e) vel[3 * ip + e] = nvel[3 * j + e];
        break;
      }
    }
    if (ip >= nmax) return -1;
  }
  return added;
}
