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

#include "aic_internal.h"
#include "panel.h"

namespace dust {

__global__ void k_geo(int n, const int *__restrict__ nver, const double *__restrict__ ver,
                      Panel *P) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  Panel g;
  panel_geo(nver[p], ver + 12 * p, g);
  P[p] = g;
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

// b(i, r) = sum_j B(i, j) s(j, r), B read once.  Workgroup (x, y) takes 64
// rows and the column range y; its four waves stride over the range's columns
// (coalesced 512-byte column segments), then the waves' and the ranges'
// partial sums are added in a fixed order.
constexpr int kRhsMax = 16;
constexpr int kRhsRows = 64;
__global__ void __launch_bounds__(256)
    k_rhs_gemv_part(int n, int nrhs, const double *__restrict__ B, long ldb,
                    const double *__restrict__ S, int per, double *part) {
  __shared__ double red[4][kRhsRows][kRhsMax];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * kRhsRows + lane;
  const int j0 = blockIdx.y * per, j1 = j0 + per < n ? j0 + per : n;
  double acc[kRhsMax];
  for (int r = 0; r < kRhsMax; ++r) acc[r] = 0.0;
  if (i < n)
    for (int j = j0 + w; j < j1; j += 4) {
      double bij = B[i + ldb * j];
#pragma unroll
      for (int r = 0; r < kRhsMax; ++r)
        if (r < nrhs) acc[r] += bij * S[(long)r * n + j];
    }
  for (int r = 0; r < kRhsMax; ++r) red[w][lane][r] = acc[r];
  __syncthreads();
  if (w == 0 && i < n)
    for (int r = 0; r < nrhs; ++r)
      part[((long)blockIdx.y * nrhs + r) * n + i] =
          ((red[0][lane][r] + red[1][lane][r]) + red[2][lane][r]) + red[3][lane][r];
}

__global__ void k_rhs_gemv_sum(int n, int nrhs, int Y, const double *__restrict__ part,
                               double *b, long ldbo) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < nrhs; ++r) {
    double s = 0.0;
    for (int y = 0; y < Y; ++y) s += part[((long)y * nrhs + r) * n + i];
    b[i + ldbo * r] = s;
  }
}

// Launch of the two kernels; `part` holds Y*nrhs*n doubles (rhs_split).
int rhs_split(int n) {
  int xb = ceil_div(n, kRhsRows);
  int Y = ceil_div(1024, xb);
  int ymax = ceil_div(n, 64);  // at least 64 columns per range
  return Y < ymax ? Y : ymax;
}

void rhs_gemv(int n, int nrhs, const double *B, long ldb, const double *S, double *part,
              double *b, long ldbo, hipStream_t st) {
  int Y = rhs_split(n);
  int per = ceil_div(n, Y);
  Y = ceil_div(n, per);
  k_rhs_gemv_part<<<dim3(ceil_div(n, kRhsRows), Y), 256, 0, st>>>(n, nrhs, B, ldb, S, per, part);
  k_rhs_gemv_sum<<<ceil_div(n, 256), 256, 0, st>>>(n, nrhs, Y, part, b, ldbo);
}

// Panel records built on the device from host vertex lists; *dP is owned
// by the caller's TmpDev.
static int geo_upload(int n, const int *n_ver, const double *ver, TmpDev<Panel> &dP,
                      hipStream_t st) {
  for (int p = 0; p < n; ++p)
    DUST_CHECK(n_ver[p] == 3 || n_ver[p] == 4, DUST_E_ARG, "panel %d has %d vertices", p, n_ver[p]);
  TmpDev<int> dn;
  TmpDev<double> dv;
  DUST_TRY(dn.alloc(n));
  DUST_TRY(dv.alloc(12 * (size_t)n));
  DUST_TRY(dP.alloc(n));
  DUST_HIP(hipMemcpyAsync(dn.p, n_ver, sizeof(int) * n, hipMemcpyHostToDevice, st));
  DUST_HIP(hipMemcpyAsync(dv.p, ver, sizeof(double) * 12 * n, hipMemcpyHostToDevice, st));
  k_geo<<<ceil_div(n, 128), 128, 0, st>>>(n, dn.p, dv.p, dP.p);
  DUST_HIP(hipStreamSynchronize(st));
  return 0;
}

static int aic_common(bool surfpan, int n, const int *n_ver, const double *ver,
                      const dust_aic_params *prm, double *A, int lda, double *B, int ldb,
                      bool dev, hipStream_t st) {
  DUST_CHECK(n >= 1 && n_ver && ver && A && prm, DUST_E_ARG, "invalid AIC arguments");
  DUST_CHECK(lda >= n && (!B || ldb >= n), DUST_E_ARG, "leading dimension smaller than n");
  TmpDev<Panel> dP;
  if (dev) {
    // n_ver and ver are device pointers
    DUST_TRY(dP.alloc(n));
    k_geo<<<ceil_div(n, 128), 128, 0, st>>>(n, n_ver, ver, dP.p);
  } else {
    DUST_TRY(geo_upload(n, n_ver, ver, dP, st));
  }
  TmpDev<double> tA, tB;
  double *dA = A, *dB = B;
  if (!dev) {
    DUST_TRY(tA.alloc((size_t)n * n));
    dA = tA.p;
    if (B) {
      DUST_TRY(tB.alloc((size_t)n * n));
      dB = tB.p;
    }
  }
  long la = dev ? lda : n, lb = dev ? ldb : n;
  TmpDev<int> dwarn;
  DUST_TRY(dwarn.alloc(1));
  DUST_HIP(hipMemsetAsync(dwarn.p, 0, sizeof(int), st));
  dim3 grid(ceil_div(n, 256), n);
  if (surfpan)
    k_aic_surfpan<<<grid, 256, 0, st>>>(n, dP.p, prm->ff_ratio_doublet, dA, la, dB, lb, dwarn.p);
  else
    k_aic_vortlatt<<<grid, 256, 0, st>>>(n, dP.p, prm->ff_ratio_doublet, prm->rankine_rad,
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
  return 0;
}

// Workspaces of the device right-hand-side entry points, reused across calls.
struct AicWs {
  DevBuf<Panel> P;
  DevBuf<double> S, U, Pt, Ptp;
};
AicWs g_aws;

void aic_release() {
  g_aws.P.release();
  g_aws.S.release();
  g_aws.U.release();
  g_aws.Pt.release();
  g_aws.Ptp.release();
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
  TmpDev<Panel> dP;
  DUST_TRY(geo_upload(n, n_ver, ver, dP, st));
  TmpDev<double> dB, dS, du, dw, db, dpart;
  DUST_TRY(dB.alloc((size_t)n * n));
  DUST_TRY(dS.alloc((size_t)n * nrhs));
  DUST_TRY(du.alloc(3 * (size_t)nrhs));
  DUST_TRY(db.alloc((size_t)n * nrhs));
  if (uvort) DUST_TRY(dw.alloc(3 * (size_t)n));
  DUST_TRY(dpart.alloc((size_t)rhs_split(n) * nrhs * n));
  DUST_HIP(hipMemcpy2D(dB.p, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, n,
                       hipMemcpyHostToDevice));
  DUST_HIP(hipMemcpy(du.p, u_inf, sizeof(double) * 3 * nrhs, hipMemcpyHostToDevice));
  if (uvort) DUST_HIP(hipMemcpy(dw.p, uvort, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  k_rhs_coef<<<ceil_div(n, 256), 256>>>(n, nrhs, dP.p, du.p, dw.p, dS.p);
  rhs_gemv(n, nrhs, dB.p, n, dS.p, dpart.p, db.p, n, nullptr);
  DUST_HIP(hipGetLastError());
  DUST_HIP(hipMemcpy2D(b, sizeof(double) * ldbo, db.p, sizeof(double) * n, sizeof(double) * n, nrhs,
                       hipMemcpyDeviceToHost));
  return 0;
}

int dust_aic_rhs_dev(int n, const int *n_ver, const double *ver, const double *B, int ldb,
                     int nrhs, const double *u_inf, const double *uvort, double *b, int ldbo,
                     double *sigma, void *stream) {
  DUST_CHECK(n >= 1 && n_ver && ver && B && u_inf && b && nrhs >= 1 && nrhs <= kRhsMax,
             DUST_E_ARG, "invalid RHS arguments (nrhs must be 1..%d)", kRhsMax);
  DUST_CHECK(ldb >= n && ldbo >= n, DUST_E_ARG, "leading dimension smaller than n");
  hipStream_t st = (hipStream_t)stream;
  DevBuf<Panel> &P = g_aws.P;
  DevBuf<double> &S = g_aws.S, &U = g_aws.U, &Pt = g_aws.Pt;
  DUST_TRY(P.reserve(n));
  DUST_TRY(U.reserve(3 * (size_t)kRhsMax));
  double *dS = sigma;
  if (!dS) {
    DUST_TRY(S.reserve((size_t)n * nrhs));
    dS = S.p;
  }
  DUST_HIP(hipMemcpyAsync(U.p, u_inf, sizeof(double) * 3 * nrhs, hipMemcpyHostToDevice, st));
  k_geo<<<ceil_div(n, 128), 128, 0, st>>>(n, n_ver, ver, P.p);
  k_rhs_coef<<<ceil_div(n, 256), 256, 0, st>>>(n, nrhs, P.p, U.p, uvort, dS);
  DUST_TRY(Pt.reserve((size_t)rhs_split(n) * nrhs * n));
  rhs_gemv(n, nrhs, B, ldb, dS, Pt.p, b, ldbo, st);
  DUST_HIP(hipGetLastError());
  // the host copy of u_inf may be reused as soon as this returns
  DUST_HIP(hipStreamSynchronize(st));
  return 0;
}

// Source part (b) of the pressure right-hand side: b_pres = B_sp src with the
// surface panels' source matrix and their bernoulli_source.
int dust_pres_rhs_src_dev(int n_sp, const double *B, int ldb, const double *src, double *b_pres,
                          void *stream) {
  DUST_CHECK(n_sp >= 1 && B && src && b_pres && ldb >= n_sp, DUST_E_ARG,
             "invalid pressure source-part arguments");
  DevBuf<double> &Pt = g_aws.Ptp;
  DUST_TRY(Pt.reserve((size_t)rhs_split(n_sp) * n_sp));
  rhs_gemv(n_sp, 1, B, ldb, src, Pt.p, b_pres, n_sp, (hipStream_t)stream);
  DUST_HIP(hipGetLastError());
  return 0;
}

int dust_pres_rhs_src(int n_sp, const double *B, int ldb, const double *src, double *b_pres) {
  DUST_CHECK(n_sp >= 1 && B && src && b_pres && ldb >= n_sp, DUST_E_ARG,
             "invalid pressure source-part arguments");
  TmpDev<double> dB, dv;
  DUST_TRY(dB.alloc((size_t)n_sp * n_sp));
  DUST_TRY(dv.alloc(2 * (size_t)n_sp));
  DUST_HIP(hipMemcpy2D(dB.p, sizeof(double) * n_sp, B, sizeof(double) * ldb, sizeof(double) * n_sp,
                       n_sp, hipMemcpyHostToDevice));
  DUST_HIP(hipMemcpy(dv.p, src, sizeof(double) * n_sp, hipMemcpyHostToDevice));
  DUST_TRY(dust_pres_rhs_src_dev(n_sp, dB.p, n_sp, dv.p, dv.p + n_sp, nullptr));
  DUST_HIP(hipMemcpy(b_pres, dv.p + n_sp, sizeof(double) * n_sp, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
