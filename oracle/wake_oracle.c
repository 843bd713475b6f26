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

/* avoid_collision (mod_wake.f90:2350-2534) for one particle: the velocity
 * is corrected element by element, in list order, against the surface panels
 * (kind 0) and vortex-lattice panels (kind 1) of the body.  Element e has
 * centre cen(3), unit normal nor(3), largest edge maxedge and velocity
 * ub(3).  dt = sim_param%dt, dtw = dt * ndt_update_wake, u_ref, rad_mult,
 * elrad_mult as in sim_param.  blthick starts at the particle's r_Vortex and,
 * as in the reference, doubles each time the vortex-lattice branch reaches
 * its thickness test. */
void ao_avoid_collision(int ne, const int *kind, const double *ecen, const double *enor,
                        const double *emaxedge, const double *eub, const double pos1[3],
                        double r_vortex, double dt, double dtw, double u_ref, double rad_mult,
                        double elrad_mult, const double vel_in[3], double vel_out[3]) {
  double vel[3] = {vel_in[0], vel_in[1], vel_in[2]};
  double blthick = r_vortex;
  const double r2d2 = sqrt(2.0) / 2.0;
  for (int ie = 0; ie < ne; ++ie) {
    if (kind[ie] != 0 && kind[ie] != 1) continue;
    const double *c = &ecen[3 * ie], *n = &enor[3 * ie], *ub = &eub[3 * ie];
    double dist1[3], relvel[3];
    for (int k = 0; k < 3; ++k) dist1[k] = pos1[k] - c[k];
    double elrad = emaxedge[ie] * r2d2;
    double check_radius = dt * u_ref * rad_mult + elrad;
    double distn = sqrt(dist1[0] * dist1[0] + dist1[1] * dist1[1] + dist1[2] * dist1[2]);
    if (!(distn < check_radius)) continue;
    for (int k = 0; k < 3; ++k) relvel[k] = vel[k] - ub[k];
    double normvel = relvel[0] * n[0] + relvel[1] * n[1] + relvel[2] * n[2];
    double dist1_nor, dt_part;
    if (kind[ie] == 0) {
      dist1_nor = dist1[0] * n[0] + dist1[1] * n[1] + dist1[2] * n[2];
      if (dist1_nor < 0.0 || normvel >= 0.0) continue;
      dt_part = fmin(fmax(dist1_nor / (-normvel), 0.0), dt);
    } else {
      double q[3] = {dist1[0] * n[0], dist1[1] * n[1], dist1[2] * n[2]};
      dist1_nor = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
      double s = dist1[0] * n[0] + dist1[1] * n[1] + dist1[2] * n[2];
      if (s * normvel >= 0.0) continue;
      dt_part = fmin(fmax(fabs(dist1_nor) / fabs(normvel), 0.0), dt);
    }
    double dist2[3];
    for (int k = 0; k < 3; ++k) dist2[k] = (pos1[k] + relvel[k] * dt_part) - c[k];
    double dist2_nor;
    if (kind[ie] == 0) {
      dist2_nor = dist2[0] * n[0] + dist2[1] * n[1] + dist2[2] * n[2];
    } else {
      double q[3] = {dist2[0] * n[0], dist2[1] * n[1], dist2[2] * n[2]};
      dist2_nor = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
      blthick = 2.0 * blthick;
    }
    if (!(dist2_nor <= blthick)) continue;
    double t2[3];
    for (int k = 0; k < 3; ++k) t2[k] = dist2[k] - n[k] * dist2_nor;
    double dist2_tan = sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
    if (!(dist2_tan < elrad_mult * elrad)) continue;
    double normvel_corr;
    if (dt_part < dt) {
      normvel_corr = -dist1_nor / dtw;
    } else {
      double dampf = dist2_nor / blthick;
      normvel_corr = normvel + (blthick - dist2_nor) / dtw * dampf;
    }
    double newvel[3];
    for (int k = 0; k < 3; ++k) newvel[k] = vel[k] + (normvel_corr - normvel) * n[k];
    normvel_corr = newvel[0] * n[0] + newvel[1] * n[1] + newvel[2] * n[2];
    double nv = vel[0] * n[0] + vel[1] * n[1] + vel[2] * n[2];
    double tanvel[3];
    for (int k = 0; k < 3; ++k) tanvel[k] = vel[k] - nv * n[k];
    double tanvel2 = tanvel[0] * tanvel[0] + tanvel[1] * tanvel[1] + tanvel[2] * tanvel[2];
    double nveldiff = nv * nv - normvel_corr * normvel_corr + tanvel2;
    if (tanvel2 > 0.0 && normvel_corr * nv >= 0.0 && nveldiff >= 0.0) {
      double f = sqrt(nveldiff) / sqrt(tanvel2);
      for (int k = 0; k < 3; ++k) tanvel[k] = tanvel[k] * f;
    }
    for (int k = 0; k < 3; ++k) vel[k] = tanvel[k] + normvel_corr * n[k];
  }
  for (int k = 0; k < 3; ++k) vel_out[k] = vel[k];
}
