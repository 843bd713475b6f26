// Evaluation (L2P + P2P) work decomposition shared by the FP64 and the tuned
// kernels of apply_multipole (src/fmm/mod_octree.f90:1139-1429).
//
// The targets of every leaf are cut into work units of at most 64 targets,
// one wavefront per unit.  A unit with nt <= 32 targets gives each target
// lpt = 64 / nt lanes (rounded down to a power of two) that split the source
// list between them, so that small leaves still fill the wavefront.
#pragma once
#include "fmm_geom.h"
#include "wind.h"

namespace dust {

constexpr int kWave = 64;
constexpr int kEvalWaves = 4;     // wavefronts per block, each on its own item

// A unit: up to 64 targets of one leaf, evaluated by nseg items.
struct EvalUnit {
  int leaf;   // index into the leaf list
  int t0;     // first target (sorted index)
  int nt;     // targets in the unit
  int lpt;    // lanes per target: a power of two, lpt * nt <= 64
  int item0;  // first item of the unit
  int nseg;   // items of the unit
  int pad0, pad1;
};

// An item: the sources [lo, hi) of the unit's P2P list (positions in the
// concatenation of its ranges), one wavefront.  Long lists are cut into
// several items so that no wavefront runs much longer than the others.
struct EvalItem {
  int unit, lo, hi, pad;
};

struct EvalArgs {
  TreeGeom G;
  int deg, nmon, use_vs, use_vd, use_tv;
  int kernel;  // DUST_KERNEL_ROSENHEAD or DUST_KERNEL_GAUSSIAN
  double nu_inf, uinf[3];
  const int *pwr, *leaves, *slot, *fstart;
  const double *mp_b, *mp_c;
  const double *px, *py, *pz, *dx, *dy, *dz, *pm, *pr;
  const int2 *rng;
  const int *cnt;
  int cap;
  const EvalUnit *units;
  const EvalItem *items;
  int n_units, n_items;
  const double *leaf_ros;  // leaf-averaged rate of strain (use_vd && use_tv)
  double near2;            // tuned form: pairs with d^2 < near2 * r^2 run in FP64
  double *part;            // per item and target: P2P velocity (3), stretching (3)
                           // and, with divf, rotu (3): pstride doubles
  int divf, pstride;       // divergence filter's rotu (use_vs && use_divfilt)
  double *vel, *str, *tv, *rot;
  WindDev wind;            // the free-stream term: variable_wind(pos, time)
  double time;
};

// Gaussian core of kernel_coeffs (the commented alternative of
// src/aeroel/mod_vortpart.f90:283-291, same expression order): velocity
// coefficient c and stretching coefficient d = (1/r) dc/dr at distance
// r = sqrt(d2) from a source of core radius rv; zero at r <= 1e-13.
__device__ inline void gauss_cd(double d2, double rv, double &c, double &d) {
  const double r = sqrt(d2);
  if (r > 1e-13) {
    const double s2p = 2.5066282746310002;  // sqrt(2 pi)
    const double e = erf(r / (1.4142135623730951 * rv));
    const double x = exp(-(r * r) / (2.0 * rv * rv));
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r;
    c = -e / r3 + 2.0 / (r2 * s2p * rv) * x;
    d = 3.0 / r5 * e + x * (-6.0 / (r4 * rv * s2p) - 2.0 / (r2 * rv * rv * rv * s2p));
  } else {
    c = 0.0;
    d = 0.0;
  }
}

// Ordering of LDS traffic inside one wavefront (no block barrier needed).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// L2P of one target at offset d from the leaf centre (mod_octree.f90:1185-1196):
// vel = sum_m b(:,m) d^m, grad(i,j) = sum_m c(i,j,m) d^m, coefficient arrays
// in LDS (b: 3 per monomial, c: 9 per monomial, column-major (i,j)).
__device__ inline void l2p_point(int deg, int nmon, int use_vs, const double *sb, const double *sc,
                                 const int *spw, double d0, double d1, double d2, double v[3],
                                 double g[9]) {
  double p0[11], p1[11], p2[11];
  p0[0] = p1[0] = p2[0] = 1.0;
  for (int e = 1; e <= deg; ++e) {
    p0[e] = p0[e - 1] * d0;
    p1[e] = p1[e - 1] * d1;
    p2[e] = p2[e - 1] * d2;
  }
  v[0] = v[1] = v[2] = 0.0;
  for (int q = 0; q < 9; ++q) g[q] = 0.0;
  for (int m = 0; m < nmon; ++m) {
    double pr = p0[spw[3 * m]] * p1[spw[3 * m + 1]] * p2[spw[3 * m + 2]];
    v[0] += sb[3 * m] * pr;
    v[1] += sb[3 * m + 1] * pr;
    v[2] += sb[3 * m + 2] * pr;
    if (use_vs)
      for (int q = 0; q < 9; ++q) g[q] += sc[9 * m + q] * pr;
  }
}

// Sum over the lanes that share a target (lanes t*lpt .. t*lpt+lpt-1,
// lpt a power of two).
__device__ inline double lane_group_sum(double v, int lpt) {
  for (int o = 1; o < lpt; o <<= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// First tile of the item's segment: position of `lo` in the range list.
__device__ inline void seg_seek(const int2 *rl, int nr, int lo, int &ri, int &off) {
  int cum = 0;
  for (ri = 0; ri < nr; ++ri) {
    int len = rl[ri].y - rl[ri].x;
    if (cum + len > lo) break;
    cum += len;
  }
  off = lo - cum;
}

}  // namespace dust
