// Direct element <-> particle terms of a DUST time step on gfx950:
//  * the element loops of apply_multipole (src/fmm/mod_octree.f90:1350-1401):
//    body panels, wake panels, wake rings and end vortices acting on the
//    vortex particles -- velocity, and with vs_elems the stretching
//    transpose(grad) alpha;
//  * get_vort_vel (src/aeroel/mod_surfpan.f90:962-977,
//    src/aeroel/mod_vortlatt.f90:676-693): particles and end vortices acting
//    on the panel centres (uvort).
//
// Kernels per element (compute_vel / compute_grad, all times 4 pi):
//   surface panel  velocity_calc_doublet * mag - velocity_calc_sou_surfpan * sigma
//                  (mod_surfpan.f90:641-658, 268-357); gradient
//                  gradient_calc_doublet * mag (the source gradient is zero,
//                  mod_surfpan.f90:362-370)
//   vortex ring    velocity_calc_doublet * mag, gradient_calc_doublet * mag
//                  (mod_doublet.f90:178-305)
//   vortex line    compute_vel_vortline / compute_grad_vortline
//                  (mod_vortline.f90:126-236)
//   particle       c (dist x dir) mag, c = -1/(|dist|^2 + r^2)^1.5
//                  (mod_vortpart.f90:139-153, 260-271)
//
// Layout: element records (a far-field block, then the geometry of
// calc_geo_data, strength, flags) are built once per call in one array, read
// by every thread with uniform loads; one thread per target point, targets in
// Morton order.  The velocity runs as a far kernel (every pair's far field,
// few registers) and a near kernel (the elements each wave flagged); the
// stretching as a kernel of its own.  get_vort_vel has few targets (panels) and
// many sources (particles): the sources are split over blockIdx.y, staged
// through LDS 256 at a time, and the per-split partial sums are added in a
// fixed order by a second kernel.
#include <cfloat>
#include <cmath>
#include <hipcub/hipcub.hpp>

#include "panel.h"

namespace dust {
namespace {

constexpr double kInv4Pi = 1.0 / (4.0 * kPi);
constexpr int kFlagDoublet = 1, kFlagSource = 2, kFlagStretch = 4;
constexpr int kSortMin = 4096;  // targets below this count are evaluated in input order
constexpr int kFillBlocks = 2048, kMinPerSplit = 32;  // element splits for few targets

// The far fields of a polygon, doublet plus source, as one formula: with
// r = x - cen, rn = r.n and the PANAIR basis t1, t2, n (t1 need not be normal
// to n on a warped panel, so the basis is not orthonormal),
//   t1 phix + t2 phiy + n pdou - sigma (t1 sx + t2 sy + n sz)
//     = (3 A m rn / r^5 + A sigma / r^3) P r - (A m / r^3) n,
//   P = t1 t1^T + t2 t2^T + n n^T  (symmetric: xx yy zz xy xz yz),
// and the two far-field switches with their thresholds squared.
struct FarCoef {
  int kind, flags;
  double cen[3], nor[3];
  double P[6];
  double am3, am, as;       // 3 A m, A m, A sigma (zero without the term)
  double td, td_lo, td_hi;  // doublet switch: ffr maxedge, its square -/+ a band
  double ts, ts_lo, ts_hi;  // source switch: 10 maxedge, likewise
};

struct FarCoef32 {  // the same in FP32 for the tuned form
  float nor[3], P[6], am3, am, as;
};

struct ElemRec {
  FarCoef f;
  FarCoef32 h;
  Panel g;  // polygons: calc_geo_data; lines: ver[0..1], evec/euni/elen[0], cen
  double sinT[4], cosT[4];  // edge directions on the PANAIR basis: euni.t2, euni.t1
  double mag, sigma;
};

// sqrt(r2) > T, as the reference decides its far-field switches (radius >
// ratio * max edge, with radius = norm2 rounded once): outside a relative
// band of 1e-14 around T^2 the answer follows from r2 and the rounded T^2
// alone (the roundings of the square root and of T^2 move it by less than
// 1e-15); inside the band the correctly rounded square root decides.
__device__ inline void switch_band(double T, double &lo, double &hi) {
  const double T2 = T * T;
  lo = T2 * (1.0 - 1e-14);
  hi = T2 * (1.0 + 1e-14);
}
__device__ inline bool beyond(double r2, double T, double lo, double hi) {
  if (r2 > hi) return true;
  if (r2 < lo) return false;
  return sqrt(r2) > T;
}

__global__ void k_elem_recs(int n, int kind, const int *__restrict__ nver,
                            const double *__restrict__ ver, const double *__restrict__ mag,
                            const double *__restrict__ sigma, int stretch, double ffr,
                            ElemRec *R) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  ElemRec r{};
  if (kind == DUST_ELEM_LINE) {
    // calc_geo_data_vortline (mod_vortline.f90:238-258)
    r.g.nv = 2;
    for (int e = 0; e < 3; ++e) {
      r.g.ver[0][e] = ver[6 * j + e];
      r.g.ver[1][e] = ver[6 * j + 3 + e];
      r.g.cen[e] = 0.5 * (r.g.ver[0][e] + r.g.ver[1][e]);
      r.g.evec[0][e] = r.g.ver[1][e] - r.g.ver[0][e];
    }
    double el = sqrt(dot3_rn(r.g.evec[0], r.g.evec[0]));
    r.g.elen[0] = el;
    r.g.maxedge = el;
    for (int e = 0; e < 3; ++e) r.g.euni[0][e] = r.g.evec[0][e] / el;
  } else {
    int nv = nver[j] == 3 ? 3 : 4;  // other counts are rejected on the host
    panel_geo(nv, ver + 12 * j, r.g);
  }
  r.mag = mag ? mag[j] : 0.0;
  r.sigma = sigma ? sigma[j] : 0.0;
  r.f.kind = kind;
  r.f.flags = (mag ? kFlagDoublet : 0) | (sigma && kind == DUST_ELEM_SURFPAN ? kFlagSource : 0) |
            (stretch ? kFlagStretch : 0);
  FarCoef &f = r.f;
  const Panel &g = r.g;
  for (int i = 0; i < 4; ++i) {
    r.sinT[i] = dot3(g.euni[i], g.t2);
    r.cosT[i] = dot3(g.euni[i], g.t1);
  }
  for (int e = 0; e < 3; ++e) {
    f.cen[e] = g.cen[e];
    f.nor[e] = g.nor[e];
  }
  const double *b[3] = {g.t1, g.t2, g.nor};
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    double acc = 0.0;
    for (int k = 0; k < 3; ++k) acc += b[k][ia[q]] * b[k][ib[q]];
    f.P[q] = acc;
  }
  const double am = (r.f.flags & kFlagDoublet) ? g.area * r.mag : 0.0;
  f.am3 = 3.0 * am;
  f.am = am;
  f.as = (r.f.flags & kFlagSource) ? g.area * r.sigma : 0.0;
  f.td = ffr * g.maxedge;
  switch_band(f.td, f.td_lo, f.td_hi);
  f.ts = 10.0 * g.maxedge;
  switch_band(f.ts, f.ts_lo, f.ts_hi);
  for (int e = 0; e < 3; ++e) r.h.nor[e] = (float)f.nor[e];
  for (int q = 0; q < 6; ++q) r.h.P[q] = (float)f.P[q];
  r.h.am3 = (float)f.am3;
  r.h.am = (float)f.am;
  r.h.as = (float)f.as;
  R[j] = r;
}

// compute_vel_vortline without the strength: the edge x1 -> x2 (unit vector
// eu, length el) with the Rankine core and the cut-off (the same expression
// as an edge of velocity_calc_doublet).
__device__ inline void line_vel(const double *x1, const double *x2, const double *eu, double el,
                                const double pos[3], double rran, double rcut, double v[3]) {
  double av[3], bv[3], hv[3], cr[3];
  for (int e = 0; e < 3; ++e) {
    av[e] = pos[e] - x1[e];
    bv[e] = pos[e] - x2[e];
  }
  double ai = dot3(av, eu);
  double R1 = sqrt(dot3(av, av)), R2 = sqrt(dot3(bv, bv));
  for (int e = 0; e < 3; ++e) hv[e] = av[e] - ai * eu[e];
  double hi = sqrt(dot3(hv, hv));
  crs3(eu, hv, cr);
  double f = 0.0;
  if (hi > el * rran)
    f = ((el - ai) / R2 + ai / R1) / (hi * hi);
  else if (R1 > el * rcut && R2 > el * rcut)
    f = ((el - ai) / R2 + ai / R1) / ((rran * el) * (rran * el));
  for (int e = 0; e < 3; ++e) v[e] = f * cr[e];
}

// transpose(grad) alpha for one edge of gradient_calc_doublet /
// compute_grad_vortline (mod_doublet.f90:254-305, mod_vortline.f90:174-236),
// accumulated into s.  With a_k = l x R_k, a = |a_1|, A = a^2 + del^2 and the
// skew matrix lx of l, the reference's endpoint term is
//   (a lx + (1/a - 3a/A) a_k a_k^T lx)(l.R) + a (a_k l^T - a_k R^T (l.R)/|R|^2)
// over A^1.5 |R|; its transpose applied to alpha is
//   (l.R)(a (alpha x l) + (1/a - 3a/A)(a_k.alpha)(a_k x l))
//   + a (a_k.alpha)(l - R (l.R)/|R|^2),
// evaluated here without forming the 3x3 matrices, with the reciprocal
// square roots of a^2, A and |R|^2 and no division.
__device__ inline void edge_gradT(const double *l, const double *x1, const double *x2,
                                  const double pos[3], double del2, const double al[3],
                                  double s[3]) {
  double R[2][3], av[2][3];
  for (int e = 0; e < 3; ++e) {
    R[0][e] = pos[e] - x1[e];
    R[1][e] = pos[e] - x2[e];
  }
  crs3(l, R[0], av[0]);
  crs3(l, R[1], av[1]);
  double a2 = dot3(av[0], av[0]);
  double ia = rsqrt64(a2);  // 1/a
  double a = a2 * ia;
  double A = a2 + del2;
  double iAh = rsqrt64(A);
  double iA15 = iAh * iAh * iAh;
  double coef = ia - 3.0 * a * (iAh * iAh);
  double axl[3];
  crs3(al, l, axl);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    double nR2 = dot3(R[t], R[t]);
    double iR = rsqrt64(nR2);
    double lR = dot3(l, R[t]);
    double fac = (t ? -iA15 : iA15) * iR;
    double ava = dot3(av[t], al);
    double avxl[3];
    crs3(av[t], l, avxl);
    double q = lR * (iR * iR);
    for (int e = 0; e < 3; ++e)
      s[e] += fac * (lR * (a * axl[e] + coef * ava * avxl[e]) + a * ava * (l[e] - R[t][e] * q));
  }
}

__device__ inline void elem_gradT(const ElemRec &E, const double pos[3], double del2,
                                  const double al[3], double s[3]) {
  const int nv = E.g.nv;
  if (nv == 2) {
    edge_gradT(E.g.euni[0], E.g.ver[0], E.g.ver[1], pos, del2, al, s);
    return;
  }
  for (int i = 0; i < nv; ++i) {
    int ip1 = (i + 1 < nv) ? i + 1 : 0;
    edge_gradT(E.g.euni[i], E.g.ver[i], E.g.ver[ip1], pos, del2, al, s);
  }
}

// The far-field test of a polygon at x: r = x - cen, r2 = |r|^2 rounded as
// the reference's norm2, and true when both far fields apply (doublet beyond
// its ratio, and the source -- if any -- beyond its own).
__device__ inline bool far_pair(const FarCoef &f, int flags, const double x[3], double rv[3],
                                double &r2) {
  for (int e = 0; e < 3; ++e) rv[e] = x[e] - f.cen[e];
  r2 = dot3_rn(rv, rv);
  const bool far_d = beyond(r2, f.td, f.td_lo, f.td_hi);  // velocity_calc_doublet switch
  const bool far_s = beyond(r2, f.ts, f.ts_lo, f.ts_hi);  // velocity_calc_sou_surfpan switch
  return far_d && (far_s || !(flags & kFlagSource));
}

// Both far fields with one reciprocal square root:
//   w = (3 A m rn / r^5 + A sigma / r^3) P r - (A m / r^3) n.
template <int TUNED>
__device__ inline void far_vel(const FarCoef &f, const FarCoef32 &h, const double rv[3], double r2,
                               double w[3]) {
  if (TUNED) {  // FP32 on the FP64 offset, FP64 sums
    const float r0 = (float)rv[0], r1 = (float)rv[1], rz = (float)rv[2];
    const float ir = __builtin_amdgcn_rsqf((float)r2);
    const float ir2 = ir * ir, ir3 = ir2 * ir;
    const float rn = r0 * h.nor[0] + r1 * h.nor[1] + rz * h.nor[2];
    const float a = (h.am3 * rn * ir2 + h.as) * ir3, b = h.am * ir3;
    const float p0 = h.P[0] * r0 + h.P[3] * r1 + h.P[4] * rz;
    const float p1 = h.P[3] * r0 + h.P[1] * r1 + h.P[5] * rz;
    const float p2 = h.P[4] * r0 + h.P[5] * r1 + h.P[2] * rz;
    w[0] = (double)(a * p0 - b * h.nor[0]);
    w[1] = (double)(a * p1 - b * h.nor[1]);
    w[2] = (double)(a * p2 - b * h.nor[2]);
  } else {
    const double ir = rsqrt64(r2);
    const double ir2 = ir * ir, ir3 = ir2 * ir;
    const double rn = dot3(rv, f.nor);
    const double a = (f.am3 * rn * ir2 + f.as) * ir3, b = f.am * ir3;
    const double p0 = f.P[0] * rv[0] + f.P[3] * rv[1] + f.P[4] * rv[2];
    const double p1 = f.P[3] * rv[0] + f.P[1] * rv[1] + f.P[5] * rv[2];
    const double p2 = f.P[4] * rv[0] + f.P[5] * rv[1] + f.P[2] * rv[2];
    w[0] = a * p0 - b * f.nor[0];
    w[1] = a * p1 - b * f.nor[1];
    w[2] = a * p2 - b * f.nor[2];
  }
}

// The full kernels of a polygon (near field of either term), times strength:
// velocity_calc_doublet (mod_doublet.f90:178-252: the ring of the edges with
// the Rankine core and the cut-off) times mag, minus velocity_calc_sou_surfpan
// (mod_surfpan.f90:268-357: edge logarithms on the projected vertices and the
// solid angle of potential_calc_doublet, mod_doublet.f90:100-176) times sigma,
// in one pass:
//  * the distances R_i to the vertices are taken once and serve the ring's
//    two edges at each vertex and the solid angle (the reference recomputes
//    them per edge and per routine, with the same expression);
//  * the ring edge factor (el - a)/R_2 + a/R_1 uses the reciprocals 1/R_i;
//  * the Rankine switch hi > el rran compares squares (the two branches
//    agree at the switch);
//  * each solid-angle term is atan2(u.(m x p), m.p), the reference's
//    atan2(sinB, cosB) without the common positive factor 1/(|m||p|), and the
//    unit vector u to the vertex is a product with 1/R_i.
// Rounding differs from the reference's order by a few ulps per term.
// TUNED = 1 (the tuned form): the same terms, the vertex distances and their
// reciprocals from the hardware reciprocal square root with two Newton steps
// and the divisions from the hardware reciprocal with two Newton steps
// (instead of IEEE square roots and divisions); every term stays FP64.  FP32
// transcendentals were measured and dropped: the four arctangents of the
// solid angle nearly cancel for points close to the panel (1e-5 relative
// velocity error in FP32), and the source edges' logarithms cancel to
// O(1/r^2) across the closed polygon for points ten edges away (6e-7).
__device__ inline double rsq_nr(double a) {
  double y = __builtin_amdgcn_rsq(a);
  y = y * (1.5 - 0.5 * a * y * y);
  return y * (1.5 - 0.5 * a * y * y);
}
__device__ inline double rcp_nr(double a) {
  double y = __builtin_amdgcn_rcp(a);
  y = fma(fma(-a, y, 1.0), y, y);
  return fma(fma(-a, y, 1.0), y, y);
}

template <int TUNED>
__device__ inline void near_vel(const ElemRec &E, const double x[3], double ffr, double rran,
                                double rcut, double w[3]) {
  const Panel &g = E.g;
  const int nv = g.nv;
  const bool dbl = E.f.flags & kFlagDoublet, src = E.f.flags & kFlagSource;
  const double rv[3] = {x[0] - g.cen[0], x[1] - g.cen[1], x[2] - g.cen[2]};
  const double radius = sqrt(dot3_rn(rv, rv));
  const bool dfar = radius > ffr * g.maxedge;   // doublet velocity and potential switch
  const bool sfar = radius > 10.0 * g.maxedge;  // source velocity switch
  const double rn = dot3(rv, g.nor);
  double vd[3] = {0.0, 0.0, 0.0};
  double R[4], iR[4];  // the offsets x - ver_i are recomputed where used (exact)
  if (!dfar && (dbl || (src && !sfar)))
    for (int i = 0; i < nv; ++i) {
      const double d[3] = {x[0] - g.ver[i][0], x[1] - g.ver[i][1], x[2] - g.ver[i][2]};
      if (TUNED) {
        const double d2 = dot3(d, d);
        iR[i] = rsq_nr(d2);
        R[i] = d2 * iR[i];
      } else {
        R[i] = sqrt(dot3(d, d));
        iR[i] = 1.0 / R[i];
      }
    }
  if (dbl) {
    if (dfar) {
      const double r2 = radius * radius, r5 = r2 * r2 * radius;
      const double rati = 3.0 * g.area * rn / r5;
      const double px = rati * dot3(rv, g.t1), py = rati * dot3(rv, g.t2);
      const double pd = g.area * (-r2 + 3.0 * rn * rn) / r5;
      for (int e = 0; e < 3; ++e) vd[e] = g.t1[e] * px + g.t2[e] * py + g.nor[e] * pd;
    } else {
      for (int i = 0; i < nv; ++i) {
        const int ip1 = (i + 1 < nv) ? i + 1 : 0;
        const double el = g.elen[i];
        const double d[3] = {x[0] - g.ver[i][0], x[1] - g.ver[i][1], x[2] - g.ver[i][2]};
        const double ai = dot3(d, g.euni[i]);
        double hv[3], cr[3];
        for (int e = 0; e < 3; ++e) hv[e] = d[e] - ai * g.euni[i][e];
        const double h2 = dot3(hv, hv);
        crs3(g.euni[i], hv, cr);
        const double rr = rran * el;
        double f = 0.0;
        if (h2 > rr * rr)
          f = TUNED ? ((el - ai) * iR[ip1] + ai * iR[i]) * rcp_nr(h2)
                    : ((el - ai) * iR[ip1] + ai * iR[i]) / h2;
        else if (R[i] > el * rcut && R[ip1] > el * rcut)
          f = TUNED ? ((el - ai) * iR[ip1] + ai * iR[i]) * rcp_nr(rr * rr)
                    : ((el - ai) * iR[ip1] + ai * iR[i]) / (rr * rr);
        for (int e = 0; e < 3; ++e) vd[e] += f * cr[e];
      }
    }
  }
  double vs[3] = {0.0, 0.0, 0.0};
  if (src) {
    double phix, phiy, pdou;
    if (sfar) {
      const double r3 = radius * radius * radius;
      phix = -g.area * dot3(rv, g.t1) / r3;
      phiy = -g.area * dot3(rv, g.t2) / r3;
      pdou = -g.area * rn / r3;
    } else {
      double Rp[4];
      for (int i = 0; i < nv; ++i) {
        double a[3];
        for (int e = 0; e < 3; ++e) a[e] = x[e] - g.verp[i][e];
        if (TUNED) {
          const double a2 = dot3(a, a);
          Rp[i] = a2 * rsq_nr(a2);
        } else {
          Rp[i] = sqrt(dot3(a, a));
        }
      }
      double sx = 0.0, sy = 0.0;
      for (int i = 0; i < nv; ++i) {
        const int ip1 = (i + 1 < nv) ? i + 1 : 0;
        const double el = g.elen[i], sr = Rp[i] + Rp[ip1];
        double sl = 0.0;
        if (sr - el >= 1e-12)
          sl = TUNED ? log((sr + el) * rcp_nr(sr - el)) : log((sr + el) / (sr - el));
        sx += E.sinT[i] * sl;
        sy -= E.cosT[i] * sl;
      }
      double dou;
      if (dfar) {
        dou = rn * g.area / (radius * radius * radius);
      } else {
        dou = 0.0;
        for (int i = 0; i < nv; ++i) {
          const int im1 = (i == 0) ? nv - 1 : i - 1;
          double u[3], ap1[3], am1[3], cr[3];
          for (int e = 0; e < 3; ++e) u[e] = (g.ver[i][e] - x[e]) * iR[i];
          const double d1 = dot3(u, g.evec[i]), d2 = dot3(u, g.evec[im1]);
          for (int e = 0; e < 3; ++e) {
            ap1[e] = g.evec[i][e] - u[e] * d1;
            am1[e] = -g.evec[im1][e] + u[e] * d2;
          }
          crs3(am1, ap1, cr);
          dou += atan2(dot3(u, cr), dot3(am1, ap1));
        }
        const double lim = (double)(nv - 2) * kPi;
        if (dou < -lim + 1.0e-5)
          dou += lim;
        else if (dou > lim - 1.0e-5)
          dou -= lim;
      }
      phix = -sx;
      phiy = -sy;
      pdou = -dou;
    }
    for (int e = 0; e < 3; ++e) vs[e] = g.t1[e] * phix + g.t2[e] * phiy + g.nor[e] * pdou;
  }
  for (int e = 0; e < 3; ++e) w[e] = vd[e] * E.mag - vs[e] * E.sigma;
}

// A line's velocity (all distances; lines have no far-field form).
__device__ inline void line_term(const ElemRec &E, const double x[3], double rran, double rcut,
                                 double w[3]) {
  w[0] = w[1] = w[2] = 0.0;
  if (E.f.flags & kFlagDoublet)
    line_vel(E.g.ver[0], E.g.ver[1], E.g.euni[0], E.g.elen[0], x, rran, rcut, w);
  for (int e = 0; e < 3; ++e) w[e] *= E.mag;
}

// The vs_elems stretching transpose(grad) alpha: one thread per target
// point, every thread walks its element range (uniform loads of the
// records).  The reference has no far-field form for it (the edge gradients
// run for every pair).  Writes slots 3..5 of the partials, or stretch.
__global__ void __launch_bounds__(256)
    k_elem_stretch(int ne, int per, const ElemRec *__restrict__ R, double del2, int n,
                   const double *__restrict__ pos, const double *__restrict__ alpha,
                   double *stretch, int acc, double *part) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
  const double al[3] = {alpha[3 * p], alpha[3 * p + 1], alpha[3 * p + 2]};
  double t[3] = {0.0, 0.0, 0.0};
  const int j0 = blockIdx.y * per;
  const int j1 = j0 + per < ne ? j0 + per : ne;
  for (int j = j0; j < j1; ++j) {
    const ElemRec &E = R[j];
    const int flags = E.f.flags;
    if ((flags & kFlagStretch) && (flags & kFlagDoublet)) {
      double sg[3] = {0.0, 0.0, 0.0};
      elem_gradT(E, x, del2, al, sg);
      for (int e = 0; e < 3; ++e) t[e] += sg[e] * E.mag;
    }
  }
  if (part) {
    double *q = part + 6 * ((long)blockIdx.y * n + p) + 3;
    for (int e = 0; e < 3; ++e) q[e] = t[e];
    return;
  }
  for (int e = 0; e < 3; ++e)
    stretch[3 * p + e] = (acc ? stretch[3 * p + e] : 0.0) + t[e] * kInv4Pi;
}

// The velocity, in two kernels, so that the far fields (most pairs) run
// with few registers and many waves per SIMD while the near-field code
// (several edges with square roots, logarithms and arctangents) keeps its
// own occupancy.  k_elem_far: lines and the far pairs; for each polygon a wave
// sets one bit of its near mask when any of its lanes is inside the switch.
// k_elem_near: per wave, the polygons of its mask, near lanes only.  Per
// target the velocity is the far terms in element order plus the near terms
// in element order (whatever the other lanes of the wave are).  Mask words:
// [split][wave][ceil(per / 64)], wave = global thread index / 64.
template <int TUNED>
__global__ void __launch_bounds__(256)
    k_elem_far(int ne, int per, const ElemRec *__restrict__ R, double rran, double rcut, int n,
               const double *__restrict__ pos, double *vel, int acc, double *part,
               unsigned long long *mask) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;  // lane 0 of a wave is live whenever the wave is
  const double x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
  double t[3] = {0.0, 0.0, 0.0};
  const int j0 = blockIdx.y * per;
  const int j1 = j0 + per < ne ? j0 + per : ne;
  const int nw = (per + 63) >> 6;
  const int wave = (blockIdx.x * blockDim.x + __builtin_amdgcn_readfirstlane(threadIdx.x)) >> 6;
  unsigned long long *mw = mask + ((long)blockIdx.y * (gridDim.x * (blockDim.x >> 6)) + wave) * nw;
  const bool lane0 = (threadIdx.x & 63) == 0;
  unsigned long long word = 0;
  for (int j = j0; j < j1; ++j) {
    const ElemRec &E = R[j];
    const FarCoef f = E.f;  // one batch of uniform loads, then one wait
    double w[3] = {0.0, 0.0, 0.0};
    if (f.kind == DUST_ELEM_LINE) {
      line_term(E, x, rran, rcut, w);
    } else {
      double rv[3], r2;
      const bool far = far_pair(f, f.flags, x, rv, r2);
      if (far) {
        if (TUNED) {
          const FarCoef32 h = E.h;
          far_vel<1>(f, h, rv, r2, w);
        } else {
          far_vel<0>(f, E.h, rv, r2, w);
        }
      }
      if (__ballot(!far)) word |= 1ull << ((j - j0) & 63);
    }
    for (int e = 0; e < 3; ++e) t[e] += w[e];
    if (((j - j0) & 63) == 63 || j == j1 - 1) {
      if (lane0) mw[(j - j0) >> 6] = word;
      word = 0;
    }
  }
  if (part) {  // one element range of several: raw partial sums
    double *q = part + 6 * ((long)blockIdx.y * n + p);
    for (int e = 0; e < 3; ++e) q[e] = t[e];
    return;
  }
  for (int e = 0; e < 3; ++e) vel[3 * p + e] = (acc ? vel[3 * p + e] : 0.0) + t[e] * kInv4Pi;
}

template <int TUNED>
__global__ void __launch_bounds__(256)
    k_elem_near(int ne, int per, const ElemRec *__restrict__ R, double ffr, double rran,
                double rcut, int n, const double *__restrict__ pos, double *vel, double *part,
                const unsigned long long *__restrict__ mask) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
  double t[3] = {0.0, 0.0, 0.0};
  const int j0 = blockIdx.y * per;
  const int j1 = j0 + per < ne ? j0 + per : ne;
  const int nw = (per + 63) >> 6, nwr = (j1 - j0 + 63) >> 6;
  const int wave = (blockIdx.x * blockDim.x + __builtin_amdgcn_readfirstlane(threadIdx.x)) >> 6;
  const unsigned long long *mw =
      mask + ((long)blockIdx.y * (gridDim.x * (blockDim.x >> 6)) + wave) * nw;
  bool any = false;
  for (int k = 0; k < nwr; ++k) {
    unsigned long long word = mw[k];
    while (word) {
      const int j = j0 + 64 * k + __builtin_ctzll(word);
      word &= word - 1;
      const ElemRec &E = R[j];
      double rv[3], r2;
      if (!far_pair(E.f, E.f.flags, x, rv, r2)) {
        double w[3];
        near_vel<TUNED>(E, x, ffr, rran, rcut, w);
        for (int e = 0; e < 3; ++e) t[e] += w[e];
        any = true;
      }
    }
  }
  if (!any) return;
  if (part) {
    double *q = part + 6 * ((long)blockIdx.y * n + p);
    for (int e = 0; e < 3; ++e) q[e] += t[e];
  } else {
    for (int e = 0; e < 3; ++e) vel[3 * p + e] += t[e] * kInv4Pi;
  }
}

// Sum of the S element-range partials of each target, in range order.
__global__ void k_elem_reduce(int n, int S, const double *__restrict__ part, double *vel,
                              double *stretch, int acc) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int nv = stretch ? 6 : 3;
  for (int y = 0; y < S; ++y) {
    const double *q = part + 6 * ((long)y * n + p);
    for (int e = 0; e < nv; ++e) v[e] += q[e];
  }
  for (int e = 0; e < 3; ++e) vel[3 * p + e] = (acc ? vel[3 * p + e] : 0.0) + v[e] * kInv4Pi;
  if (stretch)
    for (int e = 0; e < 3; ++e)
      stretch[3 * p + e] = (acc ? stretch[3 * p + e] : 0.0) + v[3 + e] * kInv4Pi;
}

// ---- Morton order of the targets -------------------------------------------

// Per-block bounding box of the targets (grid-stride), part[6 b + (lo xyz, hi xyz)].
__global__ void __launch_bounds__(256) k_bbox(int n, const double *__restrict__ pos,
                                              double *part) {
  __shared__ double sm[6][256];
  const int tid = threadIdx.x;
  double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (long p = (long)blockIdx.x * 256 + tid; p < n; p += (long)gridDim.x * 256)
    for (int e = 0; e < 3; ++e) {
      double x = pos[3 * p + e];
      v[e] = fmin(v[e], x);
      v[3 + e] = fmax(v[3 + e], x);
    }
  for (int k = 0; k < 6; ++k) sm[k][tid] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int k = 0; k < 6; ++k)
        sm[k][tid] = k < 3 ? fmin(sm[k][tid], sm[k][tid + w]) : fmax(sm[k][tid], sm[k][tid + w]);
    __syncthreads();
  }
  if (tid < 6) part[6 * blockIdx.x + tid] = sm[tid][0];
}

// Reduce nb per-block boxes into part[0..5] (one block).
__global__ void __launch_bounds__(256) k_bbox_final(int nb, double *part) {
  __shared__ double sm[6][256];
  const int tid = threadIdx.x;
  double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int b = tid; b < nb; b += 256)
    for (int k = 0; k < 6; ++k)
      v[k] = k < 3 ? fmin(v[k], part[6 * b + k]) : fmax(v[k], part[6 * b + k]);
  for (int k = 0; k < 6; ++k) sm[k][tid] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int k = 0; k < 6; ++k)
        sm[k][tid] = k < 3 ? fmin(sm[k][tid], sm[k][tid + w]) : fmax(sm[k][tid], sm[k][tid + w]);
    __syncthreads();
  }
  if (tid < 6) part[tid] = sm[tid][0];
}

__device__ inline unsigned spread10(unsigned v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton key of each target on a 1024^3 grid over the bounding box.
__global__ void k_morton(int n, const double *__restrict__ pos, const double *__restrict__ bb,
                         unsigned *key, int *idx) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  unsigned q[3];
  for (int e = 0; e < 3; ++e) {
    double w = bb[3 + e] - bb[e];
    double t = w > 0.0 ? (pos[3 * p + e] - bb[e]) * (1023.0 / w) : 0.0;
    int c = (int)t;
    q[e] = (unsigned)(c < 0 ? 0 : (c > 1023 ? 1023 : c));
  }
  key[p] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
  idx[p] = p;
}

__global__ void k_gather_targets(int n, const int *__restrict__ perm,
                                 const double *__restrict__ pos, const double *__restrict__ alpha,
                                 double *spos, double *sal) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  long p = perm[k];
  for (int e = 0; e < 3; ++e) spos[3 * (long)k + e] = pos[3 * p + e];
  if (alpha)
    for (int e = 0; e < 3; ++e) sal[3 * (long)k + e] = alpha[3 * p + e];
}

// out[perm[k]] += sorted[k] (perm is a permutation: no two threads share a target)
__global__ void k_scatter_add(int n, const int *__restrict__ perm, const double *__restrict__ svel,
                              const double *__restrict__ sst, double *vel, double *stretch) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  long p = perm[k];
  for (int e = 0; e < 3; ++e) vel[3 * p + e] += svel[3 * (long)k + e];
  if (stretch)
    for (int e = 0; e < 3; ++e) stretch[3 * p + e] += sst[3 * (long)k + e];
}

constexpr int kVT = 256;  // targets per block, one per thread
constexpr int kVS = 256;  // sources per LDS chunk

// Particle source record: x, y, z, r_vortex^2, dir*mag (3), unused.
__global__ void k_pack_particles(int np, const double *__restrict__ cen,
                                 const double *__restrict__ dir, const double *__restrict__ mag,
                                 const double *__restrict__ rv, double *out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  double m = mag[j], r = rv[j];
  double rec[8] = {cen[3 * j], cen[3 * j + 1], cen[3 * j + 2], r * r,
                   dir[3 * j] * m, dir[3 * j + 1] * m, dir[3 * j + 2] * m, 0.0};
  double4 *o = reinterpret_cast<double4 *>(out + 8 * (long)j);
  o[0] = make_double4(rec[0], rec[1], rec[2], rec[3]);
  o[1] = make_double4(rec[4], rec[5], rec[6], rec[7]);
}

// Partial uvort sums of the particles [y*per, min(np, (y+1)*per)) at the
// targets of this block: part[(y*3 + e)*m + i].
__global__ void __launch_bounds__(kVT)
    k_vort_particles(int m, const double *__restrict__ cen, int np,
                     const double *__restrict__ src, int per, double *part) {
  __shared__ double sx[kVS], sy[kVS], sz[kVS], sr[kVS], swx[kVS], swy[kVS], swz[kVS];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * kVT + tid;
  const int ic = i < m ? i : m - 1;
  const double x0 = cen[3 * ic], x1 = cen[3 * ic + 1], x2 = cen[3 * ic + 2];
  const long s0 = (long)blockIdx.y * per;
  const long s1 = s0 + per < np ? s0 + per : np;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (long base = s0; base < s1; base += kVS) {
    __syncthreads();
    long j = base + tid;
    if (j < s1) {
      const double4 *q = reinterpret_cast<const double4 *>(src + 8 * j);
      double4 u = q[0], w = q[1];
      sx[tid] = u.x; sy[tid] = u.y; sz[tid] = u.z; sr[tid] = u.w;
      swx[tid] = w.x; swy[tid] = w.y; swz[tid] = w.z;
    } else {  // zero-strength filler: contributes exactly nothing
      sx[tid] = 0.0; sy[tid] = 0.0; sz[tid] = 0.0; sr[tid] = 1.0;
      swx[tid] = 0.0; swy[tid] = 0.0; swz[tid] = 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kVS; ++k) {
      double d0 = x0 - sx[k], d1 = x1 - sy[k], d2 = x2 - sz[k];
      double R2 = fma(d0, d0, fma(d1, d1, fma(d2, d2, sr[k])));
      double ri = rsqrt64(R2);
      double c = -(ri * ri * ri);
      double w0 = swx[k], w1 = swy[k], w2 = swz[k];
      a0 = fma(c, d1 * w2 - d2 * w1, a0);
      a1 = fma(c, d2 * w0 - d0 * w2, a1);
      a2 = fma(c, d0 * w1 - d1 * w0, a2);
    }
  }
  if (i < m) {
    long o = (long)blockIdx.y * 3 * m + i;
    part[o] = a0;
    part[o + m] = a1;
    part[o + 2 * m] = a2;
  }
}

// uvort = (sum of the partials in split order + end-vortex lines) / (4 pi)
__global__ void k_vort_finish(int m, int S, const double *__restrict__ part,
                              const double *__restrict__ cen, int nl,
                              const double *__restrict__ lver, const double *__restrict__ lmag,
                              double rran, double rcut, double *uvort) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double u[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < S; ++s)
    for (int e = 0; e < 3; ++e) u[e] += part[((long)s * 3 + e) * m + i];
  if (lmag) {
    double x[3] = {cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]};
    for (int j = 0; j < nl; ++j) {
      const double *v1 = lver + 6 * j;
      double ev[3] = {v1[3] - v1[0], v1[4] - v1[1], v1[5] - v1[2]};
      double el = sqrt(dot3_rn(ev, ev));
      double eu[3] = {ev[0] / el, ev[1] / el, ev[2] / el};
      double w[3];
      line_vel(v1, v1 + 3, eu, el, x, rran, rcut, w);
      for (int e = 0; e < 3; ++e) u[e] += w[e] * lmag[j];
    }
  }
  for (int e = 0; e < 3; ++e) uvort[3 * i + e] = u[e] * kInv4Pi;
}

// Device buffers of the element terms, grown on demand (one caller at a time,
// like the rest of the library's host entry points).
struct ElemWs {
  DevBuf<ElemRec> rec;
  DevBuf<unsigned long long> nmask;  // near masks of the velocity-only path
  DevBuf<double> part, packed;
  DevBuf<unsigned> key, key_s;  // Morton order of the targets
  DevBuf<int> idx, perm;
  DevBuf<double> spos, sal, svel, sst, bb;
  DevBuf<char> tmp;
  DevBuf<double> epart;  // element-range partial sums (few targets)
  DevBuf<double> h_d;    // staging of the host entry points
  DevBuf<int> h_i;
};
ElemWs g_ws;

}  // namespace
void elem_release() {
  ElemWs &w = g_ws;
  w.rec.release(); w.nmask.release(); w.part.release(); w.packed.release();
  w.key.release(); w.key_s.release(); w.idx.release(); w.perm.release();
  w.spos.release(); w.sal.release(); w.svel.release(); w.sst.release(); w.bb.release();
  w.tmp.release(); w.epart.release(); w.h_d.release(); w.h_i.release();
}
namespace {

int check_groups(int ngroups, const dust_elem_group *groups) {
  DUST_CHECK(ngroups >= 0 && (ngroups == 0 || groups), DUST_E_ARG, "invalid element groups");
  for (int k = 0; k < ngroups; ++k) {
    const dust_elem_group &G = groups[k];
    DUST_CHECK(G.kind >= DUST_ELEM_SURFPAN && G.kind <= DUST_ELEM_LINE, DUST_E_ARG,
               "element group %d: unknown kind %d", k, G.kind);
    DUST_CHECK(G.n >= 0, DUST_E_ARG, "element group %d: negative count", k);
    if (G.n == 0) continue;
    DUST_CHECK(G.ver && (G.kind == DUST_ELEM_LINE || G.n_ver), DUST_E_ARG,
               "element group %d: missing vertices", k);
  }
  return 0;
}

int elem_apply_dev(int ngroups, const dust_elem_group *groups, const dust_elem_params *prm, int n,
                   const double *pos, const double *alpha, double *vel, double *stretch,
                   hipStream_t st) {
  DUST_TRY(check_groups(ngroups, groups));
  DUST_CHECK(prm && n >= 0 && (n == 0 || (pos && vel)), DUST_E_ARG, "invalid element arguments");
  if (!alpha) stretch = nullptr;
  long ne = 0;
  for (int k = 0; k < ngroups; ++k) ne += groups[k].n;
  DUST_CHECK(ne < (1L << 31), DUST_E_ARG, "too many elements");
  if (n == 0 || ne == 0) return 0;
  DUST_TRY(g_ws.rec.reserve(ne));
  long off = 0;
  for (int k = 0; k < ngroups; ++k) {
    const dust_elem_group &G = groups[k];
    if (G.n == 0) continue;
    k_elem_recs<<<ceil_div(G.n, 128), 128, 0, st>>>(G.n, G.kind, G.n_ver, G.ver, G.mag, G.sigma,
                                                    G.stretch, prm->ff_ratio_doublet,
                                                    g_ws.rec.p + off);
    off += G.n;
  }
  const double del2 = prm->vortex_rad * prm->vortex_rad;
  // Few targets: the element list is also split (blockIdx.y) so that the
  // launch fills the chip, and the partial sums are added in range order.
  auto launch = [&](const double *tp, const double *ta, double *vo, double *so, int acc) -> int {
    const int tb = ceil_div(n, 256);
    int S = 1;
    if (tb < kFillBlocks) {
      S = ceil_div(kFillBlocks, tb);
      int smax = ceil_div((int)ne, kMinPerSplit);
      S = S < smax ? S : smax;
    }
    int per = ceil_div((int)ne, S);
    S = ceil_div((int)ne, per);
    double *part = nullptr;
    if (S > 1) {
      DUST_TRY(g_ws.epart.reserve(6 * (size_t)S * n));
      part = g_ws.epart.p;
    }
    const bool tuned = prm->precision == DUST_ELEM_TUNED;
    const double ffr = prm->ff_ratio_doublet, rran = prm->rankine_rad, rcut = prm->cutoff_rad;
    DUST_TRY(g_ws.nmask.reserve((size_t)S * tb * 4 * ceil_div(per, 64)));
    if (tuned)
      k_elem_far<1><<<dim3(tb, S), 256, 0, st>>>((int)ne, per, g_ws.rec.p, rran, rcut, n, tp, vo,
                                                acc, part, g_ws.nmask.p);
    else
      k_elem_far<0><<<dim3(tb, S), 256, 0, st>>>((int)ne, per, g_ws.rec.p, rran, rcut, n, tp, vo,
                                                acc, part, g_ws.nmask.p);
    if (tuned)
      k_elem_near<1><<<dim3(tb, S), 256, 0, st>>>((int)ne, per, g_ws.rec.p, ffr, rran, rcut, n,
                                                  tp, vo, part, g_ws.nmask.p);
    else
      k_elem_near<0><<<dim3(tb, S), 256, 0, st>>>((int)ne, per, g_ws.rec.p, ffr, rran, rcut, n,
                                                  tp, vo, part, g_ws.nmask.p);
    if (so)
      k_elem_stretch<<<dim3(tb, S), 256, 0, st>>>((int)ne, per, g_ws.rec.p, del2, n, tp, ta, so,
                                                  acc, part);
    if (S > 1) k_elem_reduce<<<tb, 256, 0, st>>>(n, S, part, vo, so, acc);
    DUST_HIP(hipGetLastError());
    return 0;
  };
  if (n < kSortMin) return launch(pos, alpha, vel, stretch, 1);
  // Morton order: bounding box, keys, stable sort, gather, evaluate, scatter-add
  const int nb = ceil_div(n, 256) < 1024 ? ceil_div(n, 256) : 1024;
  DUST_TRY(g_ws.bb.reserve(6 * (size_t)nb));
  DUST_TRY(g_ws.key.reserve(n));
  DUST_TRY(g_ws.key_s.reserve(n));
  DUST_TRY(g_ws.idx.reserve(n));
  DUST_TRY(g_ws.perm.reserve(n));
  DUST_TRY(g_ws.spos.reserve(3 * (size_t)n));
  DUST_TRY(g_ws.svel.reserve(3 * (size_t)n));
  if (stretch) {
    DUST_TRY(g_ws.sal.reserve(3 * (size_t)n));
    DUST_TRY(g_ws.sst.reserve(3 * (size_t)n));
  }
  size_t bytes = 0;
  DUST_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, g_ws.key.p, g_ws.key_s.p,
                                              g_ws.idx.p, g_ws.perm.p, n, 0, 30, st));
  DUST_TRY(g_ws.tmp.reserve(bytes + 256));
  k_bbox<<<nb, 256, 0, st>>>(n, pos, g_ws.bb.p);
  k_bbox_final<<<1, 256, 0, st>>>(nb, g_ws.bb.p);
  k_morton<<<ceil_div(n, 256), 256, 0, st>>>(n, pos, g_ws.bb.p, g_ws.key.p, g_ws.idx.p);
  DUST_HIP(hipcub::DeviceRadixSort::SortPairs(g_ws.tmp.p, bytes, g_ws.key.p, g_ws.key_s.p,
                                              g_ws.idx.p, g_ws.perm.p, n, 0, 30, st));
  double *sal = stretch ? g_ws.sal.p : nullptr, *sst = stretch ? g_ws.sst.p : nullptr;
  k_gather_targets<<<ceil_div(n, 256), 256, 0, st>>>(n, g_ws.perm.p, pos, stretch ? alpha : nullptr,
                                                     g_ws.spos.p, sal);
  DUST_TRY(launch(g_ws.spos.p, sal, g_ws.svel.p, sst, 0));
  k_scatter_add<<<ceil_div(n, 256), 256, 0, st>>>(n, g_ws.perm.p, g_ws.svel.p, sst, vel,
                                                  stretch);
  DUST_HIP(hipGetLastError());
  return 0;
}

int vort_vel_dev(int m, const double *cen, int np, const double *pcen, const double *pdir,
                 const double *pmag, const double *prv, int nl, const double *lver,
                 const double *lmag, const dust_elem_params *prm, double *uvort, hipStream_t st) {
  DUST_CHECK(prm && m >= 0 && np >= 0 && nl >= 0, DUST_E_ARG, "invalid uvort arguments");
  DUST_CHECK(m == 0 || (cen && uvort), DUST_E_ARG, "missing centres or output");
  DUST_CHECK(np == 0 || (pcen && pdir && pmag && prv), DUST_E_ARG, "missing particle arrays");
  DUST_CHECK(nl == 0 || lver || !lmag, DUST_E_ARG, "missing end-vortex vertices");
  if (m == 0) return 0;
  const int tb = ceil_div(m, kVT);
  // enough blocks to fill the chip several times over; each split a whole
  // number of LDS chunks
  int S = 1, per = 0;
  if (np > 0) {
    S = ceil_div(4096, tb);
    int smax = ceil_div(np, kVS);
    S = S < smax ? S : smax;
    S = S > 65535 ? 65535 : S;
    per = ceil_div(ceil_div(np, S), kVS) * kVS;
    S = ceil_div(np, per);
    DUST_TRY(g_ws.packed.reserve(8 * (size_t)np));
    DUST_TRY(g_ws.part.reserve((size_t)S * 3 * m));
    k_pack_particles<<<ceil_div(np, 256), 256, 0, st>>>(np, pcen, pdir, pmag, prv, g_ws.packed.p);
    k_vort_particles<<<dim3(tb, S), kVT, 0, st>>>(m, cen, np, g_ws.packed.p, per, g_ws.part.p);
  } else {
    S = 0;
  }
  k_vort_finish<<<ceil_div(m, 256), 256, 0, st>>>(m, S, g_ws.part.p, cen, nl, lver,
                                                  nl ? lmag : nullptr, prm->rankine_rad,
                                                  prm->cutoff_rad, uvort);
  DUST_HIP(hipGetLastError());
  return 0;
}

// Host staging: copy `count` values to the device buffer at offset `off`;
// a failed copy is remembered in `err`.
template <typename T>
const T *stage(DevBuf<T> &b, size_t &off, const T *h, size_t count, hipError_t &err) {
  if (!h || count == 0) return nullptr;
  T *d = b.p + off;
  hipError_t e = hipMemcpy(d, h, sizeof(T) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) err = e;
  off += count;
  return d;
}

}  // namespace
}  // namespace dust

using namespace dust;

extern "C" {

int dust_elem_apply_dev(int ngroups, const dust_elem_group *groups, const dust_elem_params *prm,
                        int n, const double *pos, const double *alpha, double *vel,
                        double *stretch, void *stream) {
  return elem_apply_dev(ngroups, groups, prm, n, pos, alpha, vel, stretch, (hipStream_t)stream);
}

int dust_elem_apply(int ngroups, const dust_elem_group *groups, const dust_elem_params *prm,
                    int n, const double *pos, const double *alpha, double *vel, double *stretch) {
  DUST_TRY(check_groups(ngroups, groups));
  DUST_CHECK(prm && n >= 0 && (n == 0 || (pos && vel)), DUST_E_ARG, "invalid element arguments");
  for (int k = 0; k < ngroups; ++k)
    if (groups[k].kind != DUST_ELEM_LINE)
      for (int j = 0; j < groups[k].n; ++j)
        DUST_CHECK(groups[k].n_ver[j] == 3 || groups[k].n_ver[j] == 4, DUST_E_ARG,
                   "element group %d: element %d has %d vertices", k, j, groups[k].n_ver[j]);
  if (n == 0) return 0;
  const bool st = alpha && stretch;
  size_t nd = 3 * (size_t)n * (st ? 4 : 2), ni = 0;
  for (int k = 0; k < ngroups; ++k) {
    size_t g = (size_t)groups[k].n;
    nd += g * (groups[k].kind == DUST_ELEM_LINE ? 6 : 12) + 2 * g;
    ni += groups[k].kind == DUST_ELEM_LINE ? 0 : g;
  }
  DUST_TRY(g_ws.h_d.reserve(nd));
  DUST_TRY(g_ws.h_i.reserve(ni + 1));
  size_t od = 0, oi = 0;
  hipError_t err = hipSuccess;
  dust_elem_group dg[64];
  DUST_CHECK(ngroups <= 64, DUST_E_ARG, "at most 64 element groups per call");
  for (int k = 0; k < ngroups; ++k) {
    const dust_elem_group &G = groups[k];
    dg[k] = G;
    size_t g = (size_t)G.n;
    dg[k].ver = stage(g_ws.h_d, od, G.ver, g * (G.kind == DUST_ELEM_LINE ? 6 : 12), err);
    dg[k].mag = stage(g_ws.h_d, od, G.mag, g, err);
    dg[k].sigma = stage(g_ws.h_d, od, G.sigma, g, err);
    dg[k].n_ver = G.kind == DUST_ELEM_LINE ? nullptr : stage(g_ws.h_i, oi, G.n_ver, g, err);
  }
  const double *dpos = stage(g_ws.h_d, od, pos, 3 * (size_t)n, err);
  double *dvel = g_ws.h_d.p + od;
  stage(g_ws.h_d, od, (const double *)vel, 3 * (size_t)n, err);
  const double *dal = nullptr;
  double *dst = nullptr;
  if (st) {
    dal = stage(g_ws.h_d, od, alpha, 3 * (size_t)n, err);
    dst = g_ws.h_d.p + od;
    stage(g_ws.h_d, od, (const double *)stretch, 3 * (size_t)n, err);
  }
  DUST_HIP(err);
  DUST_TRY(elem_apply_dev(ngroups, dg, prm, n, dpos, dal, dvel, dst, nullptr));
  DUST_HIP(hipMemcpy(vel, dvel, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  if (st) DUST_HIP(hipMemcpy(stretch, dst, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  return 0;
}

int dust_vort_vel_dev(int m, const double *cen, int np, const double *pcen, const double *pdir,
                      const double *pmag, const double *prv, int nl, const double *lver,
                      const double *lmag, const dust_elem_params *prm, double *uvort,
                      void *stream) {
  return vort_vel_dev(m, cen, np, pcen, pdir, pmag, prv, nl, lver, lmag, prm, uvort,
                      (hipStream_t)stream);
}

int dust_vort_vel(int m, const double *cen, int np, const double *pcen, const double *pdir,
                  const double *pmag, const double *prv, int nl, const double *lver,
                  const double *lmag, const dust_elem_params *prm, double *uvort) {
  DUST_CHECK(prm && m >= 0 && np >= 0 && nl >= 0, DUST_E_ARG, "invalid uvort arguments");
  DUST_CHECK(m == 0 || (cen && uvort), DUST_E_ARG, "missing centres or output");
  DUST_CHECK(np == 0 || (pcen && pdir && pmag && prv), DUST_E_ARG, "missing particle arrays");
  DUST_CHECK(nl == 0 || lver || !lmag, DUST_E_ARG, "missing end-vortex vertices");
  if (m == 0) return 0;
  size_t nd = 6 * (size_t)m + 8 * (size_t)np + 7 * (size_t)nl;
  DUST_TRY(g_ws.h_d.reserve(nd));
  size_t od = 0;
  hipError_t err = hipSuccess;
  const double *dc = stage(g_ws.h_d, od, cen, 3 * (size_t)m, err);
  const double *dpc = stage(g_ws.h_d, od, pcen, 3 * (size_t)np, err);
  const double *dpd = stage(g_ws.h_d, od, pdir, 3 * (size_t)np, err);
  const double *dpm = stage(g_ws.h_d, od, pmag, (size_t)np, err);
  const double *dpr = stage(g_ws.h_d, od, prv, (size_t)np, err);
  const double *dlv = stage(g_ws.h_d, od, lver, 6 * (size_t)nl, err);
  const double *dlm = stage(g_ws.h_d, od, lmag, (size_t)nl, err);
  double *du = g_ws.h_d.p + od;
  DUST_HIP(err);
  DUST_TRY(vort_vel_dev(m, dc, np, dpc, dpd, dpm, dpr, nl, dlv, dlm, prm, du, nullptr));
  DUST_HIP(hipMemcpy(uvort, du, sizeof(double) * 3 * m, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
