// The panel wake of DUST on gfx950: the trailing-edge rows of initialize_wake
// and complete_wake, the panel part of update_wake, the junction of
// neighbouring trailing edges (join_te) and the particles shed from the last
// panel row.
//
// Reference behaviour (src/aeroel/mod_wake.f90):
//   initialize_wake 331-579 (rows 1 and 2 of points, pan_wake_len = 1),
//   update_wake 860-1214 (first-row intensity from the generating elements
//     884-894; convection of the points of rows 3.. 910-937; the row of end
//     points when the panel wake is full 939-955; growth of pan_wake_len
//     1133-1153; intensities shifted back one row 1177-1187),
//   complete_wake 1222-2105 (rows 1 and 2 1250-1280, join_first_panels
//     1282-1288, panel geometry 1290-1304; classical and refined shedding
//     from the last panel row 1798-1969; end vortices 2089-2103),
//   join_first_panels 2199-2313, get_joined_intensity 2315-2345.
// The RBF-interpolated shedding (interpolate_wake, 1359-1796) is not built.
//
// Layout: points and point velocities as [row][point][3] (rows 0..nmax_pan,
// the reference's 1..nmax_pan+1), intensities as [row][stripe], so that the
// panels of the reference's pan_p list (row-major over (row, stripe)) are
// contiguous for the assembly.  Every kernel is one thread per point or per
// stripe (the work is O(points x rows)); join_first_panels, which updates
// shared points pair by pair, is one thread.
#include <algorithm>
#include <vector>

#include <cstring>

#include "common.h"
#include "wake_internal.h"

struct dust_panwake_s {
  dust_panwake_params p;
  int len = 1, full = 0;
  int nend = 0, kcap = 0;
  dust::DevBuf<int> tab;          // start(2,ns) 0-based | neigh(2,ns) | neigh_o(2,ns) | gen(2,ns) | ends
  dust::DevBuf<int> joined;       // (2, nend, nmax+1)
  dust::DevBuf<double> pts, pvel; // [nmax+1][np][3]
  dust::DevBuf<double> wstart;    // [np][3] w_start_points
  dust::DevBuf<double> pend;      // [np][3] points_end
  dust::DevBuf<double> idou;      // [nmax][ns]
  dust::DevBuf<double> end_idou, last_idou;
  dust::DevBuf<double> ver;       // [nmax*ns][4][3]
  dust::DevBuf<int> nver;         // [nmax*ns]
  dust::DevBuf<double> lver, lvel;  // end vortices [ns][2][3]
  dust::DevBuf<double> rec;       // shed particles: cen(3,k) | partvec(3,k) | rv(k) | vel(3,k)
  dust::DevBuf<int> cnt;          // [0] records, [1] record overflow
};

namespace dust {
namespace {

struct PwArgs {
  int ns, np, nmax, len, full, nend, kcap;
  const int *start, *neigh, *neigh_o, *gen, *ends;
  int *joined;
  double *pts, *pvel, *wstart, *pend, *idou, *end_idou, *last_idou, *ver, *lver, *lvel;
  int *nver;
  double dt, dtw, min_vel, join_fact, kvr, vr;
  int join_te, refine, k_refine;
  double box_min[3], box_max[3];
};

__device__ inline double nrm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// Rows 1 and 2 of points (initialize_wake 512-546 with first = 1,
// complete_wake 1250-1280 with first = 0).
__global__ void k_pw_rows(PwArgs A, int first, const double *__restrict__ sp,
                          const double *__restrict__ gdir, const double *__restrict__ scal,
                          const double *__restrict__ vte, double u0, double u1, double u2) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= A.np) return;
  const double uinf[3] = {u0, u1, u2};
  double d[3], w[3], x[3];
  for (int e = 0; e < 3; ++e) {
    x[e] = sp[3 * ip + e];
    d[e] = gdir[3 * ip + e];
    w[e] = uinf[e] - vte[3 * ip + e];
    A.wstart[3 * ip + e] = x[e];
    A.pts[3 * ip + e] = x[e];
  }
  const double wn = nrm3(w), dn = nrm3(d);
  double f;
  if (wn > A.min_vel)
    f = scal[ip] * wn * A.dtw / dn;
  else
    f = first ? scal[ip] * A.min_vel * A.dtw : scal[ip] * A.min_vel * A.dtw / dn;
  for (int e = 0; e < 3; ++e) A.pts[3 * ((long)A.np + ip) + e] = x[e] + d[e] * f;
}

// join_first_panels (2199-2313), after the history of junctions has moved
// back one row (complete_wake 1282-1288; not at initialisation).  One thread:
// the pairs update shared points in order.
__global__ void k_pw_join(PwArgs A, int shift) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int ne = A.nend, R = A.nmax + 1;
  if (shift) {
    for (int r = R - 1; r >= 1; --r)
      for (int q = 0; q < 2 * ne; ++q) A.joined[(long)r * 2 * ne + q] = A.joined[(long)(r - 1) * 2 * ne + q];
    for (int q = 0; q < 2 * ne; ++q) A.joined[q] = 0;
  }
  for (int i = 0; i < ne; ++i) {
    const int iw = A.ends[i];
    const double *sp1 = A.wstart + 3 * A.start[2 * iw], *sp2 = A.wstart + 3 * A.start[2 * iw + 1];
    double t[3];
    for (int e = 0; e < 3; ++e) t[e] = sp1[e] - sp2[e];
    const double l1 = nrm3(t);
    for (int j = i + 1; j < ne; ++j) {
      const int ir = A.ends[j];
      const double *q1 = A.wstart + 3 * A.start[2 * ir], *q2 = A.wstart + 3 * A.start[2 * ir + 1];
      for (int e = 0; e < 3; ++e) t[e] = q1[e] - q2[e];
      const double l2 = nrm3(t);
      const double tol = A.join_fact * fmin(l1, l2);
      // the closest pair of end points (first of the four on ties)
      const double *a[4] = {sp1, sp1, sp2, sp2}, *b[4] = {q1, q2, q1, q2};
      int cs = 0;
      double lmin = 0.0;
      for (int c = 0; c < 4; ++c) {
        for (int e = 0; e < 3; ++e) t[e] = b[c][e] - a[c][e];
        const double l = nrm3(t);
        if (c == 0 || l < lmin) {
          lmin = l;
          cs = c;
        }
      }
      for (int e = 0; e < 3; ++e) t[e] = a[cs][e] - b[cs][e];
      if (!(nrm3(t) < tol)) continue;
      const int si = cs < 2 ? 0 : 1, sj = (cs & 1) ? 1 : 0;  // side of iw, side of ir
      const int pi = A.start[2 * iw + si], pj = A.start[2 * ir + sj];
      for (int e = 0; e < 3; ++e) {
        const double m1 = (a[cs][e] + b[cs][e]) / 2.0;
        A.pts[3 * pi + e] = m1;
        A.pts[3 * pj + e] = m1;
      }
      for (int e = 0; e < 3; ++e) {
        const double m2 = (A.pts[3 * ((long)A.np + pi) + e] + A.pts[3 * ((long)A.np + pj) + e]) / 2.0;
        A.pts[3 * ((long)A.np + pi) + e] = m2;
        A.pts[3 * ((long)A.np + pj) + e] = m2;
      }
      A.joined[2 * i + si] = ir + 1;
      A.joined[2 * j + sj] = iw + 1;
    }
  }
}

// Vertex lists of the pan_p panels (row, stripe): p1(row), p2(row),
// p2(row+1), p1(row+1).
__global__ void k_pw_geom(PwArgs A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.len * A.ns) return;
  const int row = q / A.ns, iw = q - row * A.ns;
  const int p1 = A.start[2 * iw], p2 = A.start[2 * iw + 1];
  const long r0 = (long)row * A.np, r1 = r0 + A.np;
  const long src[4] = {r0 + p1, r0 + p2, r1 + p2, r1 + p1};
  for (int v = 0; v < 4; ++v)
    for (int e = 0; e < 3; ++e) A.ver[12 * (long)q + 3 * v + e] = A.pts[3 * src[v] + e];
  A.nver[q] = 4;
}

// End vortices of a loaded full wake (load_wake 766-775): the segment
// between the stripe's points of the last row, with their velocities;
// the intensity is last_idou.
__global__ void k_pw_endvort(PwArgs A) {
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  if (iw >= A.ns) return;
  const long r = (long)A.len * A.np;
  const int p1 = A.start[2 * iw], p2 = A.start[2 * iw + 1];
  for (int e = 0; e < 3; ++e) {
    A.lver[6 * iw + e] = A.pts[3 * (r + p1) + e];
    A.lver[6 * iw + 3 + e] = A.pts[3 * (r + p2) + e];
    A.lvel[6 * iw + e] = A.pvel[3 * (r + p1) + e];
    A.lvel[6 * iw + 3 + e] = A.pvel[3 * (r + p2) + e];
  }
}

// Convection of the points (910-955): rows 3..np from the old rows 2..np-1
// with the velocities at them (vel[(k - 2) np + ip] at the 1-based row k),
// and the end points when the wake is full.
__global__ void k_pw_move(PwArgs A, const double *__restrict__ vel) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= A.np) return;
  const int np1 = A.len + 1 + (A.full ? 0 : 1);  // the reference's np
  if (A.full) {
    const double *v = vel + 3 * ((long)(A.len - 1) * A.np + ip);
    const double *x = A.pts + 3 * ((long)A.len * A.np + ip);
    for (int e = 0; e < 3; ++e) A.pend[3 * ip + e] = x[e] + v[e] * A.dtw;
  }
  for (int row = np1; row >= 3; --row) {  // 1-based, in place from the back
    const double *v = vel + 3 * ((long)(row - 3) * A.np + ip);
    const double dtr = (row != 3) ? A.dtw : A.dt;
    double *x = A.pts + 3 * ((long)(row - 1) * A.np + ip);
    const double *xo = A.pts + 3 * ((long)(row - 2) * A.np + ip);
    double *w = A.pvel + 3 * ((long)(row - 1) * A.np + ip);
    for (int e = 0; e < 3; ++e) {
      w[e] = v[e];
      x[e] = xo[e] + v[e] * dtr;
    }
  }
}

// First-row intensities from the solution (884-894), then (with the length
// of this step, new_len) end intensities and the shift back (1177-1187).
__global__ void k_pw_idou(PwArgs A, const double *__restrict__ mu, int new_len) {
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  if (iw >= A.ns) return;
  const int g1 = A.gen[2 * iw], g2 = A.gen[2 * iw + 1];
  A.idou[iw] = g2 > 0 ? mu[g1 - 1] - mu[g2 - 1] : mu[g1 - 1];
  A.end_idou[iw] = A.idou[(long)(new_len - 1) * A.ns + iw];
  for (int row = new_len; row >= 2; --row)
    A.idou[(long)(row - 1) * A.ns + iw] = A.idou[(long)(row - 2) * A.ns + iw];
}

// First-row intensities alone (update_wake 884-894), before the velocities
// of the step are evaluated with them.
__global__ void k_pw_first(PwArgs A, const double *__restrict__ mu) {
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  if (iw >= A.ns) return;
  const int g1 = A.gen[2 * iw], g2 = A.gen[2 * iw + 1];
  A.idou[iw] = g2 > 0 ? mu[g1 - 1] - mu[g2 - 1] : mu[g1 - 1];
}

// get_joined_intensity (2315-2345).
__device__ double joined_intensity(const PwArgs &A, int iw, int side) {
  int iend = 0;
  while (iend < A.nend && A.ends[iend] != iw) ++iend;
  const int ineigh = A.joined[(long)A.nmax * 2 * A.nend + 2 * iend + side] - 1;
  const double mi = A.idou[(long)(A.len - 1) * A.ns + iw];
  if (ineigh < 0) return mi;
  // normals of the first-row panels of the two stripes
  double n1[3], n2[3];
  for (int k = 0; k < 2; ++k) {
    const double *v = A.ver + 12 * (long)(k ? ineigh : iw);
    double a[3], b[3];
    for (int e = 0; e < 3; ++e) {
      a[e] = v[6 + e] - v[e];
      b[e] = v[9 + e] - v[3 + e];
    }
    double *n = k ? n2 : n1;
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
  }
  const double orient = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) >= 0.0 ? 1.0 : -1.0;
  return (mi - orient * A.idou[(long)(A.len - 1) * A.ns + ineigh]) / 2.0;
}

// Shedding from the last panel row (1798-1969) and the end vortices
// (2089-2103).  One block: thread iw forms the stripe's vortex vector and its
// particle count, a prefix sum places the records in stripe order.
__global__ void __launch_bounds__(1024) k_pw_shed(PwArgs A, double *rec, int *cnt) {
  double *rcen = rec, *rpv = rec + 3 * (long)A.kcap, *rrv = rpv + 3 * (long)A.kcap,
         *rvel = rrv + A.kcap;
  __shared__ int s_off[1024 + 1];
  const int R = A.nmax;  // 0-based row of the reference's nmax_pan+1
  int total_before = 0;
  for (int base = 0; base < A.ns; base += blockDim.x) {
    const int iw = base + threadIdx.x;
    const bool act = iw < A.ns;
    double partvec[3] = {0.0, 0.0, 0.0}, P1[3], P2[3], E1[3], E2[3], W1[3], W2[3];
    int nch = 1, nsp = 1;
    double stc = 0.0, sts = 0.0;
    if (act) {
      const int p1 = A.start[2 * iw], p2 = A.start[2 * iw + 1];
      for (int e = 0; e < 3; ++e) {
        P1[e] = A.pts[3 * ((long)R * A.np + p1) + e];
        P2[e] = A.pts[3 * ((long)R * A.np + p2) + e];
        E1[e] = A.pend[3 * p1 + e];
        E2[e] = A.pend[3 * p2 + e];
        W1[e] = A.pvel[3 * ((long)R * A.np + p1) + e];
        W2[e] = A.pvel[3 * ((long)R * A.np + p2) + e];
      }
      const double ei = A.end_idou[iw];
      double dir[3], ave;
      // left side
      for (int e = 0; e < 3; ++e) dir[e] = P1[e] - E1[e];
      const double chord = nrm3(dir);
      const int n1 = A.neigh[2 * iw];
      if (n1 > 0)
        ave = (ei - (double)A.neigh_o[2 * iw] * A.end_idou[n1 - 1]) / 2.0;
      else
        ave = A.join_te ? joined_intensity(A, iw, 0) : ei;
      for (int e = 0; e < 3; ++e) partvec[e] += dir[e] * ave;
      // right side
      for (int e = 0; e < 3; ++e) dir[e] = -P2[e] + E2[e];
      const int n2 = A.neigh[2 * iw + 1];
      if (n2 > 0)
        ave = (ei - (double)A.neigh_o[2 * iw + 1] * A.end_idou[n2 - 1]) / 2.0;
      else
        ave = A.join_te ? joined_intensity(A, iw, 1) : ei;
      for (int e = 0; e < 3; ++e) partvec[e] += dir[e] * ave;
      // end side
      for (int e = 0; e < 3; ++e) dir[e] = E1[e] - E2[e];
      const double span = nrm3(dir);
      ave = ei - A.last_idou[iw];
      for (int e = 0; e < 3; ++e) partvec[e] += dir[e] * ave;
      if (A.refine) {
        const double kr = (double)A.k_refine;
        if (chord >= span) {
          nch = (int)ceil(chord / (span / kr));
          nsp = A.k_refine;
          stc = chord / nch;
          sts = span / nsp;
        } else {
          nch = A.k_refine;
          nsp = (int)ceil(span / (chord / kr));
          stc = chord / nch;
          sts = span / nsp;
        }
      }
    }
    // exclusive prefix sum of the counts of this chunk
    const int c = act ? nch * nsp : 0;
    s_off[threadIdx.x + 1] = c;
    if (threadIdx.x == 0) s_off[0] = 0;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int t = 1; t <= (int)blockDim.x; ++t) s_off[t] += s_off[t - 1];
    __syncthreads();
    const int off = total_before + s_off[threadIdx.x];
    const int chunk_total = s_off[blockDim.x];
    if (act) {
      A.last_idou[iw] = A.end_idou[iw];
      // the end vortex of the stripe: the last row of points
      for (int e = 0; e < 3; ++e) {
        A.lver[6 * iw + e] = P1[e];
        A.lver[6 * iw + 3 + e] = P2[e];
        A.lvel[6 * iw + e] = W1[e];
        A.lvel[6 * iw + 3 + e] = W2[e];
      }
      double vel[3];
      for (int e = 0; e < 3; ++e) vel[e] = 0.5 * (W1[e] + W2[e]);
      if (!A.refine) {
        double cen[3], a[3], b[3], cr[3];
        for (int e = 0; e < 3; ++e) {
          cen[e] = (E1[e] + E2[e] + P1[e] + P2[e]) / 4.0;
          a[e] = E1[e] - P2[e];
          b[e] = E2[e] - P2[e];
        }
        cr[0] = a[1] * b[2] - a[2] * b[1];
        cr[1] = a[2] * b[0] - a[0] * b[2];
        cr[2] = a[0] * b[1] - a[1] * b[0];
        const double area = nrm3(cr);
        const double rv = A.kvr >= 1e-10 ? A.kvr * sqrt(2.0 * area) : A.vr;
        if (off < A.kcap) {
          for (int e = 0; e < 3; ++e) {
            rcen[3 * off + e] = cen[e];
            rpv[3 * off + e] = partvec[e];
            rvel[3 * off + e] = vel[e];
          }
          rrv[off] = rv;
        }
      } else {
        const double area = stc * sts;
        const double rv = A.kvr >= 1e-10 ? A.kvr * sqrt(2.0 * area) : A.vr;
        const double inv = 1.0 / (double)(nsp * nch);
        int q = off;
        for (int ic = 1; ic <= nch; ++ic)
          for (int is = 1; is <= nsp; ++is, ++q) {
            if (q >= A.kcap) continue;
            for (int e = 0; e < 3; ++e) {
              rcen[3 * q + e] = P1[e] + ((double)ic - 0.5) / (double)nch * (E1[e] - P1[e]) +
                                ((double)is - 0.5) / (double)nsp * (P2[e] - P1[e]);
              rpv[3 * q + e] = partvec[e] * inv;
              rvel[3 * q + e] = vel[e];
            }
            rrv[q] = rv;
          }
      }
    }
    total_before += chunk_total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cnt[0] = min(total_before, A.kcap);
    cnt[1] = total_before > A.kcap ? 1 : 0;
  }
}

PwArgs args(dust_panwake_s *h) {
  PwArgs A{};
  const dust_panwake_params &p = h->p;
  A.ns = p.n_stripes;
  A.np = p.n_points;
  A.nmax = p.nmax_pan;
  A.len = h->len;
  A.full = h->full;
  A.nend = h->nend;
  A.kcap = h->kcap;
  const int ns = p.n_stripes;
  A.start = h->tab.p;
  A.neigh = h->tab.p + 2 * ns;
  A.neigh_o = h->tab.p + 4 * ns;
  A.gen = h->tab.p + 6 * ns;
  A.ends = h->tab.p + 8 * ns;
  A.joined = h->joined.p;
  A.pts = h->pts.p;
  A.pvel = h->pvel.p;
  A.wstart = h->wstart.p;
  A.pend = h->pend.p;
  A.idou = h->idou.p;
  A.end_idou = h->end_idou.p;
  A.last_idou = h->last_idou.p;
  A.ver = h->ver.p;
  A.nver = h->nver.p;
  A.lver = h->lver.p;
  A.lvel = h->lvel.p;
  A.dt = p.dt;
  A.dtw = p.dt * (double)p.ndt_update_wake;
  A.min_vel = p.min_vel_at_te;
  A.join_fact = p.join_te_factor;
  A.kvr = p.k_vortex_rad;
  A.vr = p.vortex_rad;
  A.join_te = p.join_te;
  A.refine = p.refine_wake;
  A.k_refine = p.k_refine;
  for (int e = 0; e < 3; ++e) {
    A.box_min[e] = p.box_min[e];
    A.box_max[e] = p.box_max[e];
  }
  return A;
}

}  // namespace
}  // namespace dust

using namespace dust;

// Checkpoint of the device state (restart): a header of the dimensions and
// the row counters, then every state array in a fixed order.
namespace {
struct PwCkHeader {
  char magic[4];
  int version, ns, np, nmax, nend, len, full, pad;
};
template <typename T>
size_t ck_part(dust::DevBuf<T> &b, size_t count, char *out, size_t off, bool save,
               const char *in, hipError_t &err) {
  const size_t bytes = sizeof(T) * count;
  if (count) {
    hipError_t e = save ? hipMemcpy(out + off, b.p, bytes, hipMemcpyDeviceToHost)
                        : hipMemcpy(b.p, in + off, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) err = e;
  }
  return off + bytes;
}
size_t ck_arrays(dust_panwake_s *h, char *out, const char *in, bool save, hipError_t &err) {
  const size_t rows = (size_t)h->p.nmax_pan + 1, np = h->p.n_points, ns = h->p.n_stripes;
  const size_t nmax = h->p.nmax_pan, nend = h->nend;
  size_t off = sizeof(PwCkHeader);
  const bool go = out || in;
  auto part = [&](auto &buf, size_t count) {
    if (go) off = ck_part(buf, count, out, off, save, in, err);
    else off += sizeof(*buf.p) * count;
  };
  part(h->joined, 2 * nend * rows);
  part(h->pts, 3 * rows * np);
  part(h->pvel, 3 * rows * np);
  part(h->wstart, 3 * np);
  part(h->pend, 3 * np);
  part(h->idou, nmax * ns);
  part(h->end_idou, ns);
  part(h->last_idou, ns);
  part(h->ver, 12 * nmax * ns);
  part(h->nver, nmax * ns);
  part(h->lver, 6 * ns);
  part(h->lvel, 6 * ns);
  return off;
}
}  // namespace

extern "C" {

int dust_panwake_create(const dust_panwake_params *p, dust_panwake_t *out) {
  DUST_CHECK(p && out, DUST_E_ARG, "null argument");
  *out = nullptr;
  const int ns = p->n_stripes, np = p->n_points, nm = p->nmax_pan;
  DUST_CHECK(ns >= 1 && np >= 2 && nm >= 1, DUST_E_ARG, "invalid panel-wake sizes");
  DUST_CHECK(p->start_points && p->neigh && p->neigh_o && p->gen_elems, DUST_E_ARG,
             "missing panel-wake tables");
  DUST_CHECK(p->ndt_update_wake >= 1 && p->dt > 0.0, DUST_E_ARG, "invalid time step");
  DUST_CHECK(!p->refine_wake || p->k_refine >= 1, DUST_E_ARG, "k_refine must be >= 1");
  std::vector<int> tab(8 * (size_t)ns);
  std::vector<int> ends;
  for (int iw = 0; iw < ns; ++iw) {
    for (int s = 0; s < 2; ++s) {
      const int sp = p->start_points[2 * iw + s], nb = p->neigh[2 * iw + s],
                o = p->neigh_o[2 * iw + s], g = p->gen_elems[2 * iw + s];
      DUST_CHECK(sp >= 1 && sp <= np, DUST_E_ARG, "stripe %d: start point %d out of range", iw, sp);
      DUST_CHECK(nb >= 0 && nb <= ns, DUST_E_ARG, "stripe %d: neighbour %d out of range", iw, nb);
      DUST_CHECK(nb == 0 || o == 1 || o == -1, DUST_E_ARG, "stripe %d: orientation must be +-1",
                 iw);
      DUST_CHECK(g >= (s ? 0 : 1), DUST_E_ARG, "stripe %d: generating element %d", iw, g);
      tab[2 * iw + s] = sp - 1;
      tab[2 * ns + 2 * iw + s] = nb;
      tab[4 * ns + 2 * iw + s] = o;
      tab[6 * ns + 2 * iw + s] = g;
    }
    if (p->neigh[2 * iw] == 0 || p->neigh[2 * iw + 1] == 0) ends.push_back(iw);
  }
  dust_panwake_s *h = new dust_panwake_s();
  h->p = *p;
  h->p.start_points = h->p.neigh = h->p.neigh_o = h->p.gen_elems = nullptr;
  h->nend = (int)ends.size();
  // shed records: every stripe one particle, or its refined sub-particles
  h->kcap = p->refine_wake ? ns * std::max(64, 4 * p->k_refine * p->k_refine) : ns;
  tab.insert(tab.end(), ends.begin(), ends.end());
  int rc = 0;
  const size_t rows = (size_t)nm + 1;
  if ((rc = h->tab.reserve(tab.size())) || (rc = h->joined.reserve(2 * std::max(1, h->nend) * rows)) ||
      (rc = h->pts.reserve(3 * rows * np)) || (rc = h->pvel.reserve(3 * rows * np)) ||
      (rc = h->wstart.reserve(3 * (size_t)np)) || (rc = h->pend.reserve(3 * (size_t)np)) ||
      (rc = h->idou.reserve((size_t)nm * ns)) || (rc = h->end_idou.reserve(ns)) ||
      (rc = h->last_idou.reserve(ns)) || (rc = h->ver.reserve(12 * (size_t)nm * ns)) ||
      (rc = h->nver.reserve((size_t)nm * ns)) || (rc = h->lver.reserve(6 * (size_t)ns)) ||
      (rc = h->lvel.reserve(6 * (size_t)ns)) || (rc = h->rec.reserve(10 * (size_t)h->kcap)) ||
      (rc = h->cnt.reserve(2))) {
    dust_panwake_destroy(h);
    return rc;
  }
  hipError_t e = hipMemcpy(h->tab.p, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(h->joined.p, 0, sizeof(int) * h->joined.n);
  if (e == hipSuccess) e = hipMemset(h->pts.p, 0, sizeof(double) * h->pts.n);
  if (e == hipSuccess) e = hipMemset(h->pvel.p, 0, sizeof(double) * h->pvel.n);
  if (e == hipSuccess) e = hipMemset(h->pend.p, 0, sizeof(double) * h->pend.n);
  if (e == hipSuccess) e = hipMemset(h->idou.p, 0, sizeof(double) * h->idou.n);
  if (e == hipSuccess) e = hipMemset(h->end_idou.p, 0, sizeof(double) * ns);
  if (e == hipSuccess) e = hipMemset(h->last_idou.p, 0, sizeof(double) * ns);
  if (e == hipSuccess) e = hipMemset(h->lver.p, 0, sizeof(double) * 6 * ns);
  if (e == hipSuccess) e = hipMemset(h->lvel.p, 0, sizeof(double) * 6 * ns);
  if (e != hipSuccess) {
    dust_panwake_destroy(h);
    DUST_HIP(e);
  }
  *out = h;
  return 0;
}

int dust_panwake_destroy(dust_panwake_t h) {
  if (!h) return 0;
  (void)hipDeviceSynchronize();
  for (DevBuf<double> *b : {&h->pts, &h->pvel, &h->wstart, &h->pend, &h->idou, &h->end_idou,
                            &h->last_idou, &h->ver, &h->lver, &h->lvel, &h->rec})
    b->release();
  h->tab.release();
  h->joined.release();
  h->nver.release();
  h->cnt.release();
  delete h;
  return 0;
}

int dust_panwake_start_dev(dust_panwake_t h, int first, const double *start_points,
                           const double *gen_dir, const double *scaling, const double *vel_te,
                           const double u_inf[3], void *stream) {
  DUST_CHECK(h && start_points && gen_dir && scaling && vel_te && u_inf, DUST_E_ARG,
             "invalid panel-wake start arguments");
  hipStream_t st = (hipStream_t)stream;
  PwArgs A = args(h);
  k_pw_rows<<<ceil_div(A.np, 128), 128, 0, st>>>(A, first ? 1 : 0, start_points, gen_dir, scaling,
                                                 vel_te, u_inf[0], u_inf[1], u_inf[2]);
  if (h->p.join_te && h->nend > 1) k_pw_join<<<1, 64, 0, st>>>(A, first ? 0 : 1);
  k_pw_geom<<<ceil_div(A.len * A.ns, 128), 128, 0, st>>>(A);
  DUST_HIP(hipGetLastError());
  return 0;
}

int dust_panwake_panels(dust_panwake_t h, int *n_wake, int *n_impl, const int **n_ver,
                        const double **ver, const double **mag, const int **impl_ind) {
  DUST_CHECK(h && n_wake && n_impl, DUST_E_ARG, "null argument");
  *n_wake = h->len * h->p.n_stripes;
  *n_impl = h->p.n_stripes;
  if (n_ver) *n_ver = h->nver.p;
  if (ver) *ver = h->ver.p;
  if (mag) *mag = h->idou.p;
  if (impl_ind) *impl_ind = h->tab.p + 6 * h->p.n_stripes;
  return 0;
}

int dust_panwake_points(dust_panwake_t h, int *n, const double **pts) {
  DUST_CHECK(h && n && pts, DUST_E_ARG, "null argument");
  *n = h->len * h->p.n_points;
  *pts = h->pts.p + 3 * (size_t)h->p.n_points;
  return 0;
}

int dust_panwake_first_row_dev(dust_panwake_t h, const double *mu, void *stream) {
  DUST_CHECK(h && mu, DUST_E_ARG, "invalid panel-wake first-row arguments");
  PwArgs A = args(h);
  k_pw_first<<<ceil_div(A.ns, 128), 128, 0, (hipStream_t)stream>>>(A, mu);
  DUST_HIP(hipGetLastError());
  return 0;
}

int dust_panwake_update_dev(dust_panwake_t h, const double *mu, const double *vel,
                            void *stream) {
  DUST_CHECK(h && mu && vel, DUST_E_ARG, "invalid panel-wake update arguments");
  hipStream_t st = (hipStream_t)stream;
  if (h->len == h->p.nmax_pan) h->full = 1;
  PwArgs A = args(h);
  k_pw_move<<<ceil_div(A.np, 128), 128, 0, st>>>(A, vel);
  if (!h->full) h->len += 1;
  k_pw_idou<<<ceil_div(A.ns, 128), 128, 0, st>>>(A, mu, h->len);
  DUST_HIP(hipGetLastError());
  return 0;
}

int dust_panwake_shed_dev(dust_panwake_t h, int nmax, unsigned char *free_flag, double *cen,
                          double *dir, double *mag, double *r_vortex, double *vel,
                          const dust_wake_params *prm, int *n_prt, int *status, void *stream) {
  DUST_CHECK(h && prm && n_prt && status, DUST_E_ARG, "invalid panel-wake shedding arguments");
  if (!h->full) return 0;
  DUST_CHECK(nmax > 0 && free_flag && cen && dir && mag && r_vortex && vel, DUST_E_ARG,
             "missing particle arrays");
  hipStream_t st = (hipStream_t)stream;
  PwArgs A = args(h);
  const int kc = h->kcap;
  double *rec = h->rec.p;
  k_pw_shed<<<1, 1024, 0, st>>>(A, rec, h->cnt.p);
  DUST_HIP(hipGetLastError());
  double *c = rec, *pv = c + 3 * (size_t)kc, *rv = pv + 3 * (size_t)kc, *vv = rv + kc;
  return wake_insert_dev(nmax, free_flag, cen, dir, mag, r_vortex, vel, kc, h->cnt.p, c, pv, rv,
                         vv, prm, n_prt, status, st);
}

int dust_panwake_end_vortices(dust_panwake_t h, int *n, const double **ver, const double **mag,
                              const double **vel) {
  DUST_CHECK(h && n, DUST_E_ARG, "null argument");
  *n = h->full ? h->p.n_stripes : 0;
  if (ver) *ver = h->lver.p;
  if (mag) *mag = h->last_idou.p;
  if (vel) *vel = h->lvel.p;
  return 0;
}

int dust_panwake_state(dust_panwake_t h, int *len, int *full, int *nshed_max) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  if (len) *len = h->len;
  if (full) *full = h->full;
  if (nshed_max) *nshed_max = h->kcap;
  return 0;
}

int dust_panwake_get(dust_panwake_t h, double *points, double *pvel, double *idou,
                     double *end_idou, double *last_idou, double *points_end) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  DUST_HIP(hipDeviceSynchronize());
  const size_t rows = (size_t)h->p.nmax_pan + 1, np = h->p.n_points, ns = h->p.n_stripes;
  if (points) DUST_HIP(hipMemcpy(points, h->pts.p, sizeof(double) * 3 * rows * np, hipMemcpyDeviceToHost));
  if (pvel) DUST_HIP(hipMemcpy(pvel, h->pvel.p, sizeof(double) * 3 * rows * np, hipMemcpyDeviceToHost));
  if (idou) DUST_HIP(hipMemcpy(idou, h->idou.p, sizeof(double) * (rows - 1) * ns, hipMemcpyDeviceToHost));
  if (end_idou) DUST_HIP(hipMemcpy(end_idou, h->end_idou.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
  if (last_idou) DUST_HIP(hipMemcpy(last_idou, h->last_idou.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
  if (points_end) DUST_HIP(hipMemcpy(points_end, h->pend.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
  return 0;
}

int dust_panwake_checkpoint(dust_panwake_t h, void *buf, size_t *size) {
  DUST_CHECK(h && size, DUST_E_ARG, "null argument");
  hipError_t err = hipSuccess;
  const size_t need = ck_arrays(h, nullptr, nullptr, true, err);
  if (!buf) {
    *size = need;
    return 0;
  }
  DUST_CHECK(*size >= need, DUST_E_CAPACITY, "checkpoint buffer of %zu bytes, %zu needed", *size,
             need);
  DUST_HIP(hipDeviceSynchronize());
  PwCkHeader hd{{'D', 'P', 'W', 'K'}, 1, h->p.n_stripes, h->p.n_points, h->p.nmax_pan, h->nend,
                h->len, h->full, 0};
  std::memcpy(buf, &hd, sizeof(hd));
  ck_arrays(h, (char *)buf, nullptr, true, err);
  DUST_HIP(err);
  *size = need;
  return 0;
}

int dust_panwake_restore(dust_panwake_t h, const void *buf, size_t size) {
  DUST_CHECK(h && buf, DUST_E_ARG, "null argument");
  DUST_CHECK(size >= sizeof(PwCkHeader), DUST_E_ARG, "checkpoint too short");
  PwCkHeader hd;
  std::memcpy(&hd, buf, sizeof(hd));
  DUST_CHECK(std::memcmp(hd.magic, "DPWK", 4) == 0 && hd.version == 1, DUST_E_ARG,
             "not a panel-wake checkpoint");
  DUST_CHECK(hd.ns == h->p.n_stripes && hd.np == h->p.n_points && hd.nmax == h->p.nmax_pan &&
                 hd.nend == h->nend,
             DUST_E_ARG, "checkpoint of a different panel wake (%d stripes, %d points, %d rows)",
             hd.ns, hd.np, hd.nmax);
  DUST_CHECK(hd.len >= 1 && hd.len <= h->p.nmax_pan, DUST_E_ARG, "bad row count in checkpoint");
  hipError_t err = hipSuccess;
  const size_t need = ck_arrays(h, nullptr, nullptr, false, err);
  DUST_CHECK(size == need, DUST_E_ARG, "checkpoint of %zu bytes, %zu expected", size, need);
  DUST_HIP(hipDeviceSynchronize());
  ck_arrays(h, nullptr, (const char *)buf, false, err);
  DUST_HIP(err);
  h->len = hd.len;
  h->full = hd.full;
  return 0;
}

int dust_panwake_load(dust_panwake_t h, int len, const double *points, const double *pvel,
                      const double *idou, const double *last_idou, void *stream) {
  DUST_CHECK(h && points && pvel && idou && last_idou, DUST_E_ARG, "null argument");
  DUST_CHECK(len >= 1, DUST_E_ARG, "a panel wake of %d rows cannot be loaded", len);
  hipStream_t st = (hipStream_t)stream;
  const int nl = std::min(len, h->p.nmax_pan);  // pan_wake_len = min(nmax_pan, size(wvort,2))
  const size_t np = h->p.n_points, ns = h->p.n_stripes;
  const size_t rows = (size_t)nl + 1;
  // the rows the file holds beyond the kept ones are dropped (the file's
  // arrays are [len+1][np][3] and [len][ns]; only the leading rows are read)
  DUST_HIP(hipMemcpyAsync(h->pts.p, points, sizeof(double) * 3 * np * rows, hipMemcpyHostToDevice,
                          st));
  DUST_HIP(hipMemcpyAsync(h->pvel.p, pvel, sizeof(double) * 3 * np * rows, hipMemcpyHostToDevice,
                          st));
  DUST_HIP(hipMemcpyAsync(h->idou.p, idou, sizeof(double) * ns * nl, hipMemcpyHostToDevice, st));
  DUST_HIP(hipMemcpyAsync(h->last_idou.p, last_idou, sizeof(double) * ns, hipMemcpyHostToDevice,
                          st));
  h->len = nl;
  h->full = nl == h->p.nmax_pan ? 1 : 0;
  PwArgs A = args(h);
  k_pw_geom<<<ceil_div(A.len * A.ns, 128), 128, 0, st>>>(A);
  if (h->full) k_pw_endvort<<<ceil_div(A.ns, 128), 128, 0, st>>>(A);
  DUST_HIP(hipGetLastError());
  DUST_HIP(hipStreamSynchronize(st));  // the host arrays may be released on return
  return 0;
}

}  // extern "C"
