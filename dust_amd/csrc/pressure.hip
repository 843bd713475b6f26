// The Bernoulli (pressure) integral equation of a DUST time step on gfx950:
// assemble_pressure_sys (src/linsys/mod_pressure_equation.f90:139-236), the
// rotational part (c) of the right-hand side at the surface-panel centres:
//   (c.1) vortex particles (:171-187)
//         + sum_p  d.(dir_p x vel_p) mag_p / (|d|^2 + r_p^2)^(3/2),  d = cen_p - x
//   (c.2) end-vortex lines with a strength (:191-203)
//         - sum_l  mag_l/2 e_l.(d_1 x v_1 / |d_1|^3 + d_2 x v_2 / |d_2|^3)
//   (c.3) wake panels, edge by edge (:204-225)
//         - sum_w sum_is  mag_w/2 e_is.(d_is x v_is / |d_is|^3 + d_nx x v_nx / |d_nx|^3)
// with d_k = ver_k - x, e the edge vector (ver_next - ver_is; ver_2 - ver_1
// for a line) and v_k the velocity of vertex k.  Actuator-disk rings (c.4) and
// the viscous part (d) are left as TODOs by the reference and are absent here.
//
// Layout as get_vort_vel: few targets (panel centres), many sources
// (particles).  The particles are packed once as 64-byte records (position,
// r^2, w = mag dir x vel), split over blockIdx.y, staged through LDS 256 at a
// time; the per-split partial sums are added in split order by a second
// kernel, which also walks the lines and the wake panels (few) per target.
#include <cmath>

#include "panel.h"

namespace dust {
namespace {

constexpr int kPT = 256;  // targets per block
constexpr int kPS = 256;  // sources per LDS chunk

__global__ void k_pres_pack(int np, const double *__restrict__ cen,
                            const double *__restrict__ dir, const double *__restrict__ mag,
                            const double *__restrict__ rv, const double *__restrict__ vel,
                            double *out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  double c[3];
  crs3(dir + 3 * j, vel + 3 * j, c);
  const double m = mag[j], r = rv[j];
  double4 *o = reinterpret_cast<double4 *>(out + 8 * (long)j);
  o[0] = make_double4(cen[3 * j], cen[3 * j + 1], cen[3 * j + 2], r * r);
  o[1] = make_double4(c[0] * m, c[1] * m, c[2] * m, 0.0);
}

// Partial (c.1) sums of the particles [y*per, min(np, (y+1)*per)) at the
// targets of this block: part[y*m + i].
__global__ void __launch_bounds__(kPT)
    k_pres_particles(int m, const double *__restrict__ cen, int np,
                     const double *__restrict__ src, int per, double *part) {
  __shared__ double sx[kPS], sy[kPS], sz[kPS], sr[kPS], swx[kPS], swy[kPS], swz[kPS];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * kPT + tid;
  const int ic = i < m ? i : m - 1;
  const double x0 = cen[3 * ic], x1 = cen[3 * ic + 1], x2 = cen[3 * ic + 2];
  const long s0 = (long)blockIdx.y * per;
  const long s1 = s0 + per < np ? s0 + per : np;
  double a = 0.0;
  for (long base = s0; base < s1; base += kPS) {
    __syncthreads();
    const long j = base + tid;
    if (j < s1) {
      const double4 *q = reinterpret_cast<const double4 *>(src + 8 * j);
      const double4 u = q[0], w = q[1];
      sx[tid] = u.x; sy[tid] = u.y; sz[tid] = u.z; sr[tid] = u.w;
      swx[tid] = w.x; swy[tid] = w.y; swz[tid] = w.z;
    } else {  // zero-strength filler: contributes exactly nothing
      sx[tid] = 0.0; sy[tid] = 0.0; sz[tid] = 0.0; sr[tid] = 1.0;
      swx[tid] = 0.0; swy[tid] = 0.0; swz[tid] = 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kPS; ++k) {
      const double d0 = sx[k] - x0, d1 = sy[k] - x1, d2 = sz[k] - x2;
      const double R2 = fma(d0, d0, fma(d1, d1, fma(d2, d2, sr[k])));
      const double ri = rsqrt64(R2);
      const double s = fma(d0, swx[k], fma(d1, swy[k], d2 * swz[k]));
      a = fma(s, ri * ri * ri, a);
    }
  }
  if (i < m) part[(long)blockIdx.y * m + i] = a;
}

// One edge term of (c.2) / (c.3): e.(d1 x v1 / |d1|^3 + d2 x v2 / |d2|^3).
__device__ inline double edge_term(const double *p1, const double *p2, const double *v1,
                                   const double *v2, const double x[3]) {
  double d1[3], d2[3], e[3], c1[3], c2[3];
  for (int k = 0; k < 3; ++k) {
    d1[k] = p1[k] - x[k];
    d2[k] = p2[k] - x[k];
    e[k] = p2[k] - p1[k];
  }
  crs3(d1, v1, c1);
  crs3(d2, v2, c2);
  const double n1 = sqrt(dot3(d1, d1)), n2 = sqrt(dot3(d2, d2));
  const double q1 = n1 * n1 * n1, q2 = n2 * n2 * n2;
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s += e[k] * (c1[k] / q1 + c2[k] / q2);
  return s;
}

// b_pres += (sum of the particle partials in split order) - lines - wake
// panels, per target.
__global__ void k_pres_finish(int m, int S, const double *__restrict__ part,
                              const double *__restrict__ cen, int nl,
                              const double *__restrict__ lver, const double *__restrict__ lmag,
                              const double *__restrict__ lvel, int nw,
                              const int *__restrict__ wnver, const double *__restrict__ wver,
                              const double *__restrict__ wmag, const double *__restrict__ wvel,
                              double *b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double acc = b[i];
  double sp = 0.0;
  for (int s = 0; s < S; ++s) sp += part[(long)s * m + i];
  acc += sp;
  const double x[3] = {cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]};
  if (lmag)
    for (int l = 0; l < nl; ++l) {
      const double *v = lver + 6 * l, *u = lvel + 6 * l;
      acc -= 0.5 * lmag[l] * edge_term(v, v + 3, u, u + 3, x);
    }
  for (int w = 0; w < nw; ++w) {
    const int nv = wnver[w] == 3 ? 3 : 4;
    const double *v = wver + 12 * w, *u = wvel + 12 * w;
    for (int is = 0; is < nv; ++is) {
      const int nx = is + 1 < nv ? is + 1 : 0;
      acc -= 0.5 * wmag[w] * edge_term(v + 3 * is, v + 3 * nx, u + 3 * is, u + 3 * nx, x);
    }
  }
  b[i] = acc;
}

struct PresWs {
  DevBuf<double> packed, part;
  DevBuf<double> h_d;  // staging of the host entry point
  DevBuf<int> h_i;
};
PresWs g_pws;

}  // namespace
void pres_release() {
  g_pws.packed.release();
  g_pws.part.release();
  g_pws.h_d.release();
  g_pws.h_i.release();
}
namespace {

int pres_rhs_rot_dev(int m, const double *cen, int np, const double *pcen, const double *pdir,
                     const double *pmag, const double *prv, const double *pvel, int nl,
                     const double *lver, const double *lmag, const double *lvel, int nw,
                     const int *wnver, const double *wver, const double *wmag,
                     const double *wvel, double *b_pres, hipStream_t st) {
  DUST_CHECK(m >= 0 && np >= 0 && nl >= 0 && nw >= 0, DUST_E_ARG,
             "invalid pressure right-hand-side sizes");
  DUST_CHECK(m == 0 || (cen && b_pres), DUST_E_ARG, "missing centres or b_pres");
  DUST_CHECK(np == 0 || (pcen && pdir && pmag && prv && pvel), DUST_E_ARG,
             "missing particle arrays");
  DUST_CHECK(nl == 0 || !lmag || (lver && lvel), DUST_E_ARG, "missing end-vortex arrays");
  DUST_CHECK(nw == 0 || (wnver && wver && wmag && wvel), DUST_E_ARG, "missing wake-panel arrays");
  for (int w = 0; w < nw; ++w)
    DUST_CHECK(wnver[w] == 3 || wnver[w] == 4, DUST_E_ARG, "wake panel %d has %d vertices", w,
               wnver[w]);
  if (m == 0) return 0;
  const int tb = ceil_div(m, kPT);
  int S = 0, per = 0;
  if (np > 0) {
    S = ceil_div(4096, tb);  // fill the chip several times over
    const int smax = ceil_div(np, kPS);
    S = S < smax ? S : smax;
    S = S > 65535 ? 65535 : S;
    per = ceil_div(ceil_div(np, S), kPS) * kPS;  // whole LDS chunks per split
    S = ceil_div(np, per);
    DUST_TRY(g_pws.packed.reserve(8 * (size_t)np));
    DUST_TRY(g_pws.part.reserve((size_t)S * m));
    k_pres_pack<<<ceil_div(np, 256), 256, 0, st>>>(np, pcen, pdir, pmag, prv, pvel,
                                                   g_pws.packed.p);
    k_pres_particles<<<dim3(tb, S), kPT, 0, st>>>(m, cen, np, g_pws.packed.p, per,
                                                  g_pws.part.p);
  }
  k_pres_finish<<<ceil_div(m, 256), 256, 0, st>>>(m, S, g_pws.part.p, cen, nl, lver,
                                                  nl ? lmag : nullptr, lvel, nw, wnver, wver,
                                                  wmag, wvel, b_pres);
  DUST_HIP(hipGetLastError());
  return 0;
}

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

// press_normvel_der (mod_pressure_equation.f90:396-462), one thread per
// surface panel, the edges in the reference's order; the R products as
// matmul forms them (row i of R times the vector).
__global__ void k_normvel_der(int n, const double *__restrict__ nor, const double *__restrict__ ub,
                              const double *__restrict__ sv, const double *__restrict__ svo,
                              const double *__restrict__ st, const int *__restrict__ nv,
                              const int *__restrict__ ng, const double *__restrict__ R, double dt,
                              double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *ni = nor + 3 * i, *u = sv + 3 * i, *b = ub + 3 * i, *Ri = R + 9 * (long)i;
  const double dun = (ni[0] * (b[0] - svo[3 * i]) + ni[1] * (b[1] - svo[3 * i + 1]) +
                      ni[2] * (b[2] - svo[3 * i + 2])) / dt;
  double g[3] = {0.0, 0.0, 0.0};
  double div = 0.0;
  const int ne = min(max(nv[i], 0), 4);
  // GradS_Un (a neighbour index outside -1..n-1 adds nothing, as -2 does)
  for (int e = 0; e < ne; ++e) {
    const int j = ng[4 * i + e];
    if (j < -1 || j >= n) continue;
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = j >= 0 ? sv[3 * j + k] - u[k] : -2.0 * u[k];
    const double dn = ni[0] * d[0] + ni[1] * d[1] + ni[2] * d[2];
    const double *s = st + 12 * (long)i + 3 * e;
    double v[3] = {s[0] * dn, s[1] * dn, s[2] * dn};
    for (int r = 0; r < 3; ++r) g[r] += Ri[r] * v[0] + Ri[3 + r] * v[1] + Ri[6 + r] * v[2];
  }
  // DivS_U
  for (int e = 0; e < ne; ++e) {
    const int j = ng[4 * i + e];
    if (j < -1 || j >= n) continue;
    const double *s = st + 12 * (long)i + 3 * e;
    double q = 0.0;
    for (int r = 0; r < 3; ++r) {
      const double rs = Ri[r] * s[0] + Ri[3 + r] * s[1] + Ri[6 + r] * s[2];
      q += rs * (j >= 0 ? sv[3 * j + r] - u[r] : -2.0 * u[r]);
    }
    div += q;
  }
  out[i] = dun - (g[0] * b[0] + g[1] * b[1] + g[2] * b[2]) +
           div * (b[0] * ni[0] + b[1] * ni[1] + b[2] * ni[2]);
}

static int normvel_der_dev(int n, const double *nor, const double *ub, const double *sv,
                           const double *svo, const double *st, const int *nv, const int *ng,
                           const double *R, double dt, double *out, hipStream_t stream) {
  DUST_CHECK(n >= 0, DUST_E_ARG, "negative panel count");
  DUST_CHECK(dt != 0.0, DUST_E_ARG, "dt must not be zero");
  if (n == 0) return 0;
  DUST_CHECK(nor && ub && sv && svo && st && nv && ng && R && out, DUST_E_ARG, "null argument");
  k_normvel_der<<<ceil_div(n, 128), 128, 0, stream>>>(n, nor, ub, sv, svo, st, nv, ng, R, dt, out);
  DUST_HIP(hipGetLastError());
  return 0;
}

extern "C" {

int dust_pres_rhs_rot_dev(int m, const double *cen, int np, const double *pcen,
                          const double *pdir, const double *pmag, const double *prv,
                          const double *pvel, int nl, const double *lver, const double *lmag,
                          const double *lvel, int nw, const int *wnver, const double *wver,
                          const double *wmag, const double *wvel, double *b_pres,
                          void *stream) {
  return pres_rhs_rot_dev(m, cen, np, pcen, pdir, pmag, prv, pvel, nl, lver, lmag, lvel, nw,
                          wnver, wver, wmag, wvel, b_pres, (hipStream_t)stream);
}

int dust_pres_rhs_rot(int m, const double *cen, int np, const double *pcen, const double *pdir,
                      const double *pmag, const double *prv, const double *pvel, int nl,
                      const double *lver, const double *lmag, const double *lvel, int nw,
                      const int *wnver, const double *wver, const double *wmag,
                      const double *wvel, double *b_pres) {
  DUST_CHECK(m >= 0 && np >= 0 && nl >= 0 && nw >= 0, DUST_E_ARG,
             "invalid pressure right-hand-side sizes");
  DUST_CHECK(m == 0 || (cen && b_pres), DUST_E_ARG, "missing centres or b_pres");
  DUST_CHECK(np == 0 || (pcen && pdir && pmag && prv && pvel), DUST_E_ARG,
             "missing particle arrays");
  DUST_CHECK(nl == 0 || !lmag || (lver && lvel), DUST_E_ARG, "missing end-vortex arrays");
  DUST_CHECK(nw == 0 || (wnver && wver && wmag && wvel), DUST_E_ARG, "missing wake-panel arrays");
  for (int w = 0; w < nw; ++w)
    DUST_CHECK(wnver[w] == 3 || wnver[w] == 4, DUST_E_ARG, "wake panel %d has %d vertices", w,
               wnver[w]);
  if (m == 0) return 0;
  const bool lines = nl > 0 && lmag;
  size_t nd = 4 * (size_t)m + 11 * (size_t)np + (lines ? 13 * (size_t)nl : 0) + 25 * (size_t)nw;
  DUST_TRY(g_pws.h_d.reserve(nd));
  DUST_TRY(g_pws.h_i.reserve((size_t)nw + 1));
  size_t od = 0, oi = 0;
  hipError_t err = hipSuccess;
  const double *dc = stage(g_pws.h_d, od, cen, 3 * (size_t)m, err);
  const double *dpc = stage(g_pws.h_d, od, pcen, 3 * (size_t)np, err);
  const double *dpd = stage(g_pws.h_d, od, pdir, 3 * (size_t)np, err);
  const double *dpm = stage(g_pws.h_d, od, pmag, (size_t)np, err);
  const double *dpr = stage(g_pws.h_d, od, prv, (size_t)np, err);
  const double *dpv = stage(g_pws.h_d, od, pvel, 3 * (size_t)np, err);
  const double *dlv = lines ? stage(g_pws.h_d, od, lver, 6 * (size_t)nl, err) : nullptr;
  const double *dlm = lines ? stage(g_pws.h_d, od, lmag, (size_t)nl, err) : nullptr;
  const double *dlu = lines ? stage(g_pws.h_d, od, lvel, 6 * (size_t)nl, err) : nullptr;
  const int *dwn = stage(g_pws.h_i, oi, wnver, (size_t)nw, err);
  const double *dwv = stage(g_pws.h_d, od, wver, 12 * (size_t)nw, err);
  const double *dwm = stage(g_pws.h_d, od, wmag, (size_t)nw, err);
  const double *dwu = stage(g_pws.h_d, od, wvel, 12 * (size_t)nw, err);
  double *db = g_pws.h_d.p + od;
  stage(g_pws.h_d, od, (const double *)b_pres, (size_t)m, err);
  DUST_HIP(err);
  DUST_TRY(pres_rhs_rot_dev(m, dc, np, dpc, dpd, dpm, dpr, dpv, lines ? nl : 0, dlv, dlm, dlu, nw,
                            dwn, dwv, dwm, dwu, db, nullptr));
  DUST_HIP(hipMemcpy(b_pres, db, sizeof(double) * m, hipMemcpyDeviceToHost));
  return 0;
}

int dust_press_normvel_der_dev(int n_sp, const double *nor, const double *ub,
                               const double *surf_vel, const double *surf_vel_old,
                               const double *stencil, const int *n_ver, const int *neigh,
                               const double *R, double dt, double *bernoulli_source,
                               void *stream) {
  return normvel_der_dev(n_sp, nor, ub, surf_vel, surf_vel_old, stencil, n_ver, neigh, R, dt,
                         bernoulli_source, (hipStream_t)stream);
}

int dust_press_normvel_der(int n_sp, const double *nor, const double *ub,
                           const double *surf_vel, const double *surf_vel_old,
                           const double *stencil, const int *n_ver, const int *neigh,
                           const double *R, double dt, double *bernoulli_source) {
  DUST_CHECK(n_sp >= 0, DUST_E_ARG, "negative panel count");
  DUST_CHECK(dt != 0.0, DUST_E_ARG, "dt must not be zero");
  if (n_sp == 0) return 0;
  DUST_CHECK(nor && ub && surf_vel && surf_vel_old && stencil && n_ver && neigh && R &&
                 bernoulli_source,
             DUST_E_ARG, "null argument");
  for (int i = 0; i < n_sp; ++i) {
    DUST_CHECK(n_ver[i] == 3 || n_ver[i] == 4, DUST_E_ARG, "panel %d has %d vertices", i,
               n_ver[i]);
    for (int e = 0; e < n_ver[i]; ++e)
      DUST_CHECK(neigh[4 * i + e] >= -2 && neigh[4 * i + e] < n_sp, DUST_E_ARG,
                 "panel %d: neighbour %d out of range", i, neigh[4 * i + e]);
  }
  const size_t n = (size_t)n_sp;
  const size_t nd = 3 * n * 4 + 12 * n + 9 * n + n;
  DUST_TRY(g_pws.h_d.reserve(nd));
  DUST_TRY(g_pws.h_i.reserve(5 * n));
  size_t od = 0, oi = 0;
  hipError_t err = hipSuccess;
  const double *dn = stage(g_pws.h_d, od, nor, 3 * n, err);
  const double *du = stage(g_pws.h_d, od, ub, 3 * n, err);
  const double *ds = stage(g_pws.h_d, od, surf_vel, 3 * n, err);
  const double *dso = stage(g_pws.h_d, od, surf_vel_old, 3 * n, err);
  const double *dst = stage(g_pws.h_d, od, stencil, 12 * n, err);
  const double *dR = stage(g_pws.h_d, od, R, 9 * n, err);
  const int *dnv = stage(g_pws.h_i, oi, n_ver, n, err);
  const int *dng = stage(g_pws.h_i, oi, neigh, 4 * n, err);
  double *dout = g_pws.h_d.p + od;
  DUST_HIP(err);
  DUST_TRY(normvel_der_dev(n_sp, dn, du, ds, dso, dst, dnv, dng, dR, dt, dout, nullptr));
  DUST_HIP(hipMemcpy(bernoulli_source, dout, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
