// Panel geometry and the constant-strength panel kernels, shared by the AIC
// assembly (aic.hip) and the element <-> particle terms (elem.hip).
//
// Reference behaviour: calc_geo_data_surfpan (src/aeroel/mod_surfpan.f90:890-957),
// calc_geo_data_vortlatt (src/aeroel/mod_vortlatt.f90:601-665),
// potential_calc_doublet (src/aeroel/mod_doublet.f90:113-168),
// potential_calc_sou_surfpan (mod_surfpan.f90:173-258),
// velocity_calc_doublet (mod_doublet.f90:178-248).
#pragma once
#include <cmath>

#include "common.h"

namespace dust {

// Panel record in device memory, one struct per panel (read through uniform
// loads by the column-per-block assembly kernels).
struct Panel {
  double cen[3], nor[3], t1[3], t2[3];
  double area, maxedge;
  double ver[4][3], evec[4][3], euni[4][3], verp[4][3];
  double elen[4];
  int nv, pad;
};

__device__ inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
// The same sum rounded term by term, as plain C on the host evaluates it (no
// fused multiply-add).  Used for the lengths that decide a branch (far-field
// switches, edge lengths, core radii): on regular meshes many distances sit
// exactly on a switch (a centre 10 edge lengths away), and a fused rounding
// would pick the other formula there.
__device__ inline double dot3_rn(const double *a, const double *b) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ inline void crs3(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// calc_geo_data (surface panel and vortex ring share these fields)
__device__ inline void panel_geo(int nv, const double *__restrict__ vp, Panel &g) {
  g.nv = nv;
  g.pad = 0;
  for (int i = 0; i < 4; ++i)
    for (int e = 0; e < 3; ++e) g.ver[i][e] = (i < nv) ? vp[3 * i + e] : 0.0;
  for (int e = 0; e < 3; ++e) {
    double s = 0.0;
    for (int i = 0; i < nv; ++i) s += g.ver[i][e];
    g.cen[e] = s / (double)nv;
  }
  double a[3], b[3], nr[3];
  for (int e = 0; e < 3; ++e) {
    if (nv == 4) {
      a[e] = g.ver[2][e] - g.ver[0][e];
      b[e] = g.ver[3][e] - g.ver[1][e];
    } else {
      a[e] = g.ver[2][e] - g.ver[1][e];
      b[e] = g.ver[0][e] - g.ver[1][e];
    }
  }
  crs3(a, b, nr);
  double nn = sqrt(dot3(nr, nr));
  g.area = 0.5 * nn;
  for (int e = 0; e < 3; ++e) g.nor[e] = nr[e] / nn;
  double tl[3];
  for (int e = 0; e < 3; ++e) tl[e] = 0.5 * (g.ver[nv - 1][e] + g.ver[0][e]) - g.cen[e];
  double tn = sqrt(dot3(tl, tl));
  for (int e = 0; e < 3; ++e) g.t1[e] = tl[e] / tn;
  crs3(g.nor, g.t1, g.t2);
  g.maxedge = 0.0;
  for (int i = 0; i < 4; ++i) {
    int nx = (i + 1 < nv) ? i + 1 : 0;
    for (int e = 0; e < 3; ++e) g.evec[i][e] = (i < nv) ? g.ver[nx][e] - g.ver[i][e] : 0.0;
    g.elen[i] = (i < nv) ? sqrt(dot3_rn(g.evec[i], g.evec[i])) : 0.0;
    for (int e = 0; e < 3; ++e) g.euni[i][e] = (i < nv) ? g.evec[i][e] / g.elen[i] : 0.0;
    if (i < nv && g.elen[i] > g.maxedge) g.maxedge = g.elen[i];
    double d[3];
    for (int e = 0; e < 3; ++e) d[e] = g.ver[i][e] - g.cen[e];
    double z = dot3(d, g.nor);
    for (int e = 0; e < 3; ++e) g.verp[i][e] = (i < nv) ? g.ver[i][e] - g.nor[e] * z : 0.0;
  }
}

// 1/sqrt(x) in FP64: the hardware estimate and two Newton steps.
__device__ inline double rsqrt64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// potential_calc_doublet: solid angle of panel g seen from pos.
__device__ inline double pot_doublet(const Panel &g, const double pos[3], double ffr) {
  double rv[3] = {pos[0] - g.cen[0], pos[1] - g.cen[1], pos[2] - g.cen[2]};
  double radius = sqrt(dot3_rn(rv, rv));
  double zQ = dot3(rv, g.nor);
  if (radius > ffr * g.maxedge) return zQ * g.area / (radius * radius * radius);
  double dou = 0.0;
  const int nv = g.nv;
  for (int i1 = 0; i1 < nv; ++i1) {
    int im1 = (i1 == 0) ? nv - 1 : i1 - 1;
    double ei[3] = {g.ver[i1][0] - pos[0], g.ver[i1][1] - pos[1], g.ver[i1][2] - pos[2]};
    double en = sqrt(dot3(ei, ei));
    for (int e = 0; e < 3; ++e) ei[e] /= en;
    double d1 = dot3(ei, g.evec[i1]), d2 = dot3(ei, g.evec[im1]);
    double ap1[3], am1[3];
    for (int e = 0; e < 3; ++e) {
      ap1[e] = g.evec[i1][e] - ei[e] * d1;
      am1[e] = -g.evec[im1][e] + ei[e] * d2;
    }
    double nrm = sqrt(dot3(ap1, ap1)) * sqrt(dot3(am1, am1));
    double cr[3];
    crs3(am1, ap1, cr);
    double sinB = dot3(ei, cr) / nrm;
    double cosB = dot3(am1, ap1) / nrm;
    dou += atan2(sinB, cosB);
  }
  double lim = (double)(nv - 2) * kPi;
  if (dou < -lim + 1.0e-5)
    dou += lim;
  else if (dou > lim - 1.0e-5)
    dou -= lim;
  return dou;
}

// potential_calc_sou_surfpan (the routine's local far-field ratio is 10)
__device__ inline double pot_source(const Panel &g, double dou, const double pos[3], int *warn) {
  double rv[3] = {pos[0] - g.cen[0], pos[1] - g.cen[1], pos[2] - g.cen[2]};
  double radius = sqrt(dot3_rn(rv, rv));
  if (radius > 10.0 * g.maxedge) return g.area / radius;
  double zQ = dot3(rv, g.nor);
  double Qp[3] = {pos[0] - zQ * g.nor[0], pos[1] - zQ * g.nor[1], pos[2] - zQ * g.nor[2]};
  double sou = 0.0;
  const int nv = g.nv;
  for (int i1 = 0; i1 < nv; ++i1) {
    int ip1 = (i1 + 1 < nv) ? i1 + 1 : 0;
    double a[3], b[3];
    for (int e = 0; e < 3; ++e) {
      a[e] = pos[e] - g.verp[i1][e];
      b[e] = pos[e] - g.verp[ip1][e];
    }
    double R1 = sqrt(dot3(a, a)), R2 = sqrt(dot3(b, b));
    double el = g.elen[i1];
    double den = R1 + R2 - el;
    if (den < 1e-6) {
      *warn += 1;
      for (int e = 0; e < 3; ++e) {
        a[e] = pos[e] - g.ver[i1][e];
        b[e] = pos[e] - g.ver[ip1][e];
      }
      R1 = sqrt(dot3(a, a));
      R2 = sqrt(dot3(b, b));
      den = R1 + R2 - el;
    }
    double sl = log((R1 + R2 + el) / den);
    double q[3] = {Qp[0] - g.verp[i1][0], Qp[1] - g.verp[i1][1], Qp[2] - g.verp[i1][2]};
    double cr[3];
    crs3(q, g.evec[i1], cr);
    sou += -dot3(cr, g.nor) / el * sl;
  }
  return sou - zQ * dou;
}

// velocity_calc_doublet (vortex ring of intensity 4 pi, Rankine-like core)
__device__ inline void vel_doublet(const Panel &g, const double pos[3], double ffr, double rran,
                                   double rcut, double v[3]) {
  double rv[3] = {pos[0] - g.cen[0], pos[1] - g.cen[1], pos[2] - g.cen[2]};
  double radius = sqrt(dot3_rn(rv, rv));
  v[0] = v[1] = v[2] = 0.0;
  if (radius > ffr * g.maxedge) {
    double r2 = radius * radius, r5 = r2 * r2 * radius, rn = dot3(rv, g.nor);
    double rati = 3.0 * g.area * rn / r5;
    double phix = rati * dot3(rv, g.t1), phiy = rati * dot3(rv, g.t2);
    double pdou = g.area * (-r2 + 3.0 * rn * rn) / r5;
    for (int e = 0; e < 3; ++e) v[e] = g.t1[e] * phix + g.t2[e] * phiy + g.nor[e] * pdou;
    return;
  }
  const int nv = g.nv;
  for (int i1 = 0; i1 < nv; ++i1) {
    int ip1 = (i1 + 1 < nv) ? i1 + 1 : 0;
    double el = g.elen[i1];
    double av[3], bv[3];
    for (int e = 0; e < 3; ++e) {
      av[e] = pos[e] - g.ver[i1][e];
      bv[e] = pos[e] - g.ver[ip1][e];
    }
    double ai = dot3(av, g.euni[i1]);
    double R1 = sqrt(dot3(av, av)), R2 = sqrt(dot3(bv, bv));
    double hv[3];
    for (int e = 0; e < 3; ++e) hv[e] = av[e] - ai * g.euni[i1][e];
    double hi = sqrt(dot3(hv, hv));
    double cr[3];
    crs3(g.euni[i1], hv, cr);
    if (hi > el * rran) {
      double f = ((el - ai) / R2 + ai / R1) / (hi * hi);
      for (int e = 0; e < 3; ++e) v[e] += f * cr[e];
    } else if (R1 > el * rcut && R2 > el * rcut) {
      double rr = rran * el;
      double f = ((el - ai) / R2 + ai / R1) / (rr * rr);
      for (int e = 0; e < 3; ++e) v[e] += f * cr[e];
    }
  }
}

// velocity_calc_sou_surfpan (mod_surfpan.f90:268-357): the routine's own
// far-field ratio is the constant 10; its near field takes the doublet
// potential with the doublet far-field ratio ffr.  The PANAIR edge cosines
// cosTi = edge_uni.t1, sinTi = edge_uni.t2 of calc_geo_data_surfpan are formed
// here.
__device__ inline void vel_source(const Panel &g, const double pos[3], double ffr, double v[3]) {
  const double rv[3] = {pos[0] - g.cen[0], pos[1] - g.cen[1], pos[2] - g.cen[2]};
  const double radius = sqrt(dot3_rn(rv, rv));
  double phix, phiy, pdou;
  if (radius > 10.0 * g.maxedge) {
    const double r3 = radius * radius * radius;
    phix = -g.area * dot3(rv, g.t1) / r3;
    phiy = -g.area * dot3(rv, g.t2) / r3;
    pdou = -g.area * dot3(rv, g.nor) / r3;
  } else {
    phix = 0.0;
    phiy = 0.0;
    const int nv = g.nv;
    for (int i1 = 0; i1 < nv; ++i1) {
      const int ip1 = (i1 + 1 < nv) ? i1 + 1 : 0;
      double a[3], b[3];
      for (int e = 0; e < 3; ++e) {
        a[e] = pos[e] - g.verp[i1][e];
        b[e] = pos[e] - g.verp[ip1][e];
      }
      const double R1 = sqrt(dot3(a, a)), R2 = sqrt(dot3(b, b)), el = g.elen[i1];
      const double sl = (R1 + R2 - el < 1e-12) ? 0.0 : log((R1 + R2 + el) / (R1 + R2 - el));
      phix += dot3(g.euni[i1], g.t2) * sl;
      phiy -= dot3(g.euni[i1], g.t1) * sl;
    }
    pdou = -pot_doublet(g, pos, ffr);
    phix = -phix;
    phiy = -phiy;
  }
  for (int e = 0; e < 3; ++e) v[e] = g.t1[e] * phix + g.t2[e] * phiy + g.nor[e] * pdou;
}

}  // namespace dust
