// Dense AIC assembly on gfx950: panel geometry, constant-strength doublet and
// source potentials (surface panels), vortex-ring normal velocity (vortex
// lattice), and the static right-hand side.
//
// Reference behaviour: calc_geo_data_surfpan (src/aeroel/mod_surfpan.f90:890-957),
// calc_geo_data_vortlatt (src/aeroel/mod_vortlatt.f90:601-665),
// potential_calc_doublet (src/aeroel/mod_doublet.f90:113-168),
// potential_calc_sou_surfpan (mod_surfpan.f90:173-258),
// velocity_calc_doublet (mod_doublet.f90:178-248),
// compute_pot_surfpan (mod_surfpan.f90:579-601), compute_psi_vortlatt
// (mod_vortlatt.f90:365-391), build_row_static_* and the static RHS loops
// (mod_surfpan.f90:400-404, 448-481; mod_vortlatt.f90:158-164, 201-237).
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
__device__ inline void crs3(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// calc_geo_data (surface panel and vortex ring share these fields)
__global__ void k_geo(int n, const int *__restrict__ nver, const double *__restrict__ ver,
                      Panel *P) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  Panel g;
  int nv = nver[p];
  g.nv = nv;
  g.pad = 0;
  for (int i = 0; i < 4; ++i)
    for (int e = 0; e < 3; ++e) g.ver[i][e] = (i < nv) ? ver[12 * p + 3 * i + e] : 0.0;
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
    g.elen[i] = (i < nv) ? sqrt(dot3(g.evec[i], g.evec[i])) : 0.0;
    for (int e = 0; e < 3; ++e) g.euni[i][e] = (i < nv) ? g.evec[i][e] / g.elen[i] : 0.0;
    if (i < nv && g.elen[i] > g.maxedge) g.maxedge = g.elen[i];
    double d[3];
    for (int e = 0; e < 3; ++e) d[e] = g.ver[i][e] - g.cen[e];
    double z = dot3(d, g.nor);
    for (int e = 0; e < 3; ++e) g.verp[i][e] = (i < nv) ? g.ver[i][e] - g.nor[e] * z : 0.0;
  }
  P[p] = g;
}

// potential_calc_doublet: solid angle of panel g seen from pos.
__device__ inline double pot_doublet(const Panel &g, const double pos[3], double ffr) {
  double rv[3] = {pos[0] - g.cen[0], pos[1] - g.cen[1], pos[2] - g.cen[2]};
  double radius = sqrt(dot3(rv, rv));
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
  double radius = sqrt(dot3(rv, rv));
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
  double radius = sqrt(dot3(rv, rv));
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

// One block column j (blockIdx.y), threads over rows i: coalesced stores of
// the column-major A(:,j) and B(:,j); panel j is read with uniform loads.
__global__ void __launch_bounds__(256) k_aic_surfpan(int n, const Panel *__restrict__ P,
                                                     double ffr, double *A, long lda, double *B,
                                                     long ldb, int *warn) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= n) return;
  const Panel &gj = P[j];
  double pos[3] = {P[i].cen[0], P[i].cen[1], P[i].cen[2]};
  double dou = (i != j) ? pot_doublet(gj, pos, ffr) : -2.0 * kPi;
  A[i + lda * j] = -dou;
  if (B) {
    int w = 0;
    B[i + ldb * j] = pot_source(gj, dou, pos, &w);
    if (w) atomicAdd(warn, w);
  }
}

__global__ void __launch_bounds__(256) k_aic_vortlatt(int n, const Panel *__restrict__ P,
                                                      double ffr, double rran, double rcut,
                                                      double *A, long lda, double *B, long ldb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= n) return;
  double pos[3] = {P[i].cen[0], P[i].cen[1], P[i].cen[2]};
  double v[3];
  vel_doublet(P[j], pos, ffr, rran, rcut, v);
  A[i + lda * j] = dot3(v, P[i].nor);
  if (B) B[i + ldb * j] = (i == j) ? 4.0 * kPi : 0.0;
}

// s(j, r) = nor_j . (-u_inf(:,r) - uvort(:,j))
__global__ void k_rhs_coef(int n, int nrhs, const Panel *__restrict__ P,
                           const double *__restrict__ uinf, const double *__restrict__ uvort,
                           double *S) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int r = 0; r < nrhs; ++r) {
    double w[3];
    for (int e = 0; e < 3; ++e) w[e] = -uinf[3 * r + e] - (uvort ? uvort[3 * j + e] : 0.0);
    S[(long)r * n + j] = dot3(P[j].nor, w);
  }
}

// b(i, r) = sum_j B(i, j) s(j, r), one thread per row, B read once.
constexpr int kRhsMax = 16;
__global__ void k_rhs_gemv(int n, int nrhs, const double *__restrict__ B, long ldb,
                           const double *__restrict__ S, double *b, long ldbo) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc[kRhsMax];
  for (int r = 0; r < kRhsMax; ++r) acc[r] = 0.0;
  for (int j = 0; j < n; ++j) {
    double bij = B[i + ldb * j];
#pragma unroll
    for (int r = 0; r < kRhsMax; ++r)
      if (r < nrhs) acc[r] += bij * S[(long)r * n + j];
  }
  for (int r = 0; r < nrhs; ++r) b[i + ldbo * r] = acc[r];
}

static int geo_upload(int n, const int *n_ver, const double *ver, Panel **dP, hipStream_t st) {
  for (int p = 0; p < n; ++p)
    DUST_CHECK(n_ver[p] == 3 || n_ver[p] == 4, DUST_E_ARG, "panel %d has %d vertices", p, n_ver[p]);
  int *dn = nullptr;
  double *dv = nullptr;
  DUST_HIP(hipMalloc(&dn, sizeof(int) * n));
  DUST_HIP(hipMalloc(&dv, sizeof(double) * 12 * n));
  DUST_HIP(hipMalloc(dP, sizeof(Panel) * n));
  DUST_HIP(hipMemcpyAsync(dn, n_ver, sizeof(int) * n, hipMemcpyHostToDevice, st));
  DUST_HIP(hipMemcpyAsync(dv, ver, sizeof(double) * 12 * n, hipMemcpyHostToDevice, st));
  k_geo<<<ceil_div(n, 128), 128, 0, st>>>(n, dn, dv, *dP);
  DUST_HIP(hipStreamSynchronize(st));
  (void)hipFree(dn);
  (void)hipFree(dv);
  return 0;
}

static int aic_common(bool surfpan, int n, const int *n_ver, const double *ver,
                      const dust_aic_params *prm, double *A, int lda, double *B, int ldb,
                      bool dev, hipStream_t st) {
  DUST_CHECK(n >= 1 && n_ver && ver && A && prm, DUST_E_ARG, "invalid AIC arguments");
  DUST_CHECK(lda >= n && (!B || ldb >= n), DUST_E_ARG, "leading dimension smaller than n");
  Panel *dP = nullptr;
  int *dn = nullptr;
  double *dv = nullptr;
  if (dev) {
    // n_ver and ver are device pointers
    DUST_HIP(hipMalloc(&dP, sizeof(Panel) * n));
    k_geo<<<ceil_div(n, 128), 128, 0, st>>>(n, n_ver, ver, dP);
  } else {
    DUST_TRY(geo_upload(n, n_ver, ver, &dP, st));
  }
  (void)dn;
  (void)dv;
  double *dA = A, *dB = B;
  if (!dev) {
    DUST_HIP(hipMalloc(&dA, sizeof(double) * (size_t)n * n));
    if (B) DUST_HIP(hipMalloc(&dB, sizeof(double) * (size_t)n * n));
  }
  long la = dev ? lda : n, lb = dev ? ldb : n;
  int *dwarn = nullptr;
  DUST_HIP(hipMalloc(&dwarn, sizeof(int)));
  DUST_HIP(hipMemsetAsync(dwarn, 0, sizeof(int), st));
  dim3 grid(ceil_div(n, 256), n);
  if (surfpan)
    k_aic_surfpan<<<grid, 256, 0, st>>>(n, dP, prm->ff_ratio_doublet, dA, la, dB, lb, dwarn);
  else
    k_aic_vortlatt<<<grid, 256, 0, st>>>(n, dP, prm->ff_ratio_doublet, prm->rankine_rad,
                                         prm->cutoff_rad, dA, la, dB, lb);
  DUST_HIP(hipGetLastError());
  if (!dev) {
    DUST_HIP(hipMemcpy2DAsync(A, sizeof(double) * lda, dA, sizeof(double) * n, sizeof(double) * n, n,
                              hipMemcpyDeviceToHost, st));
    if (B)
      DUST_HIP(hipMemcpy2DAsync(B, sizeof(double) * ldb, dB, sizeof(double) * n,
                                sizeof(double) * n, n, hipMemcpyDeviceToHost, st));
  }
  DUST_HIP(hipStreamSynchronize(st));
  if (!dev) {
    (void)hipFree(dA);
    if (B) (void)hipFree(dB);
  }
  (void)hipFree(dP);
  (void)hipFree(dwarn);
  return 0;
}

}  // namespace dust

using namespace dust;

extern "C" {

int dust_aic_surfpan(int n, const int *n_ver, const double *ver, const dust_aic_params *prm,
                     double *A, int lda, double *B, int ldb) {
  return aic_common(true, n, n_ver, ver, prm, A, lda, B, ldb, false, nullptr);
}

int dust_aic_surfpan_dev(int n, const int *n_ver, const double *ver, const dust_aic_params *prm,
                         double *A, int lda, double *B, int ldb, void *stream) {
  return aic_common(true, n, n_ver, ver, prm, A, lda, B, ldb, true, (hipStream_t)stream);
}

int dust_aic_vortlatt(int n, const int *n_ver, const double *ver, const dust_aic_params *prm,
                      double *A, int lda, double *B, int ldb) {
  return aic_common(false, n, n_ver, ver, prm, A, lda, B, ldb, false, nullptr);
}

int dust_aic_rhs(int n, const int *n_ver, const double *ver, const double *B, int ldb, int nrhs,
                 const double *u_inf, const double *uvort, double *b, int ldbo) {
  DUST_CHECK(n >= 1 && B && u_inf && b && nrhs >= 1 && nrhs <= kRhsMax, DUST_E_ARG,
             "invalid RHS arguments (nrhs must be 1..%d)", kRhsMax);
  DUST_CHECK(ldb >= n && ldbo >= n, DUST_E_ARG, "leading dimension smaller than n");
  hipStream_t st = nullptr;
  Panel *dP = nullptr;
  DUST_TRY(geo_upload(n, n_ver, ver, &dP, st));
  double *dB = nullptr, *dS = nullptr, *du = nullptr, *dw = nullptr, *db = nullptr;
  DUST_HIP(hipMalloc(&dB, sizeof(double) * (size_t)n * n));
  DUST_HIP(hipMalloc(&dS, sizeof(double) * (size_t)n * nrhs));
  DUST_HIP(hipMalloc(&du, sizeof(double) * 3 * nrhs));
  DUST_HIP(hipMalloc(&db, sizeof(double) * (size_t)n * nrhs));
  if (uvort) DUST_HIP(hipMalloc(&dw, sizeof(double) * 3 * n));
  DUST_HIP(hipMemcpy2D(dB, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, n,
                       hipMemcpyHostToDevice));
  DUST_HIP(hipMemcpy(du, u_inf, sizeof(double) * 3 * nrhs, hipMemcpyHostToDevice));
  if (uvort) DUST_HIP(hipMemcpy(dw, uvort, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  k_rhs_coef<<<ceil_div(n, 256), 256>>>(n, nrhs, dP, du, dw, dS);
  k_rhs_gemv<<<ceil_div(n, 128), 128>>>(n, nrhs, dB, n, dS, db, n);
  DUST_HIP(hipMemcpy2D(b, sizeof(double) * ldbo, db, sizeof(double) * n, sizeof(double) * n, nrhs,
                       hipMemcpyDeviceToHost));
  (void)hipFree(dP); (void)hipFree(dB); (void)hipFree(dS); (void)hipFree(du); (void)hipFree(db);
  if (dw) (void)hipFree(dw);
  return 0;
}

}  // extern "C"
