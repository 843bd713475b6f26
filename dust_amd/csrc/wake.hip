// Device-resident vortex-particle pool of the wake: the per-step particle
// update of complete_wake, the live-particle list of prepare_wake and the
// insertion of newly shed particles into free slots.
//
// Reference behaviour (src/aeroel/mod_wake.f90):
//   complete_wake particle loop 1305-1355: pos = cen + vel dt; inside the
//     particle box -> cen = pos and (use_vs or use_vd) optional divergence
//     filter stretch -= filt_eta (dir mag - rotu mag/|rotu|), then explicit
//     Euler alpha = dir mag + stretch dt, dir = alpha/|alpha| if |alpha| /= 0
//     (mag is left as it is); outside -> the slot is freed, n_prt -= 1;
//   prepare_wake 826-835: part_p = the non-free slots in slot order;
//   complete_wake shedding 1888-1967 (and the ring branch after it): every
//     new particle inside the box takes the lowest free slot, mag = |partvec|,
//     dir = partvec/mag (partvec itself when mag <= 1e-13), n_prt += 1; no
//     free slot left is an error.
//
// The pool is SoA in HBM (cen(3,nmax), dir(3,nmax), mag, r_vortex, vel,
// stretch, rotu, one free byte per slot).  The update is one thread per slot
// (HBM-bound: about 170 bytes per live particle); the list and the insertion
// use hipcub stream compaction of the slot flags.
#include <hipcub/hipcub.hpp>

#include "panel.h"

namespace dust {
namespace {

// x + v dt with the product rounded before the sum, as plain C on the host
// evaluates it: the box test that frees a slot then decides alike.
__device__ inline double step_rn(double x, double v, double dt) {
#pragma clang fp contract(off)
  return x + v * dt;
}

__global__ void __launch_bounds__(256)
    k_wake_update(int nmax, unsigned char *freef, double *cen, double *dir,
                  const double *__restrict__ mag, const double *__restrict__ vel,
                  double *stretch, const double *__restrict__ rotu, dust_wake_params P,
                  int *n_prt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int freed = 0;
  if (i < nmax && !freef[i]) {
    const double dt = P.dt;
    double pos[3];
    bool inside = true;
    for (int e = 0; e < 3; ++e) {
      pos[e] = step_rn(cen[3 * (long)i + e], vel[3 * (long)i + e], dt);
      inside = inside && pos[e] >= P.box_min[e] && pos[e] <= P.box_max[e];
    }
    if (inside) {
      for (int e = 0; e < 3; ++e) cen[3 * (long)i + e] = pos[e];
      if (P.use_vs || P.use_vd) {
        const double m = mag[i];
        double d[3], s[3];
        for (int e = 0; e < 3; ++e) {
          d[e] = dir[3 * (long)i + e];
          s[e] = stretch[3 * (long)i + e];
        }
        if (P.use_divfilt) {
          double r[3] = {rotu[3 * (long)i], rotu[3 * (long)i + 1], rotu[3 * (long)i + 2]};
          double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
          for (int e = 0; e < 3; ++e) {
            s[e] = s[e] - P.filt_eta * (d[e] * m - r[e] * m / rn);
            stretch[3 * (long)i + e] = s[e];
          }
        }
        double a[3];
        for (int e = 0; e < 3; ++e) a[e] = d[e] * m + s[e] * dt;
        double an = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        if (an != 0.0)
          for (int e = 0; e < 3; ++e) dir[3 * (long)i + e] = a[e] / an;
      }
    } else {
      freef[i] = 1;
      freed = 1;
    }
  }
  // one atomic per wave with a freed slot
  unsigned long long b = __ballot(freed);
  if (b && (threadIdx.x & 63) == __builtin_ctzll(b)) atomicSub(n_prt, __popcll(b));
}

// live -> contiguous arrays: pcen/pdir/pmag/prv (and palpha = dir mag) for
// the FMM and the element terms
__global__ void k_wake_gather(int n, const int *__restrict__ live, const double *__restrict__ cen,
                              const double *__restrict__ dir, const double *__restrict__ mag,
                              const double *__restrict__ rv, double *pcen, double *pdir,
                              double *pmag, double *prv, double *palpha) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  long i = live[k];
  double m = mag[i];
  for (int e = 0; e < 3; ++e) {
    double d = dir[3 * i + e];
    pcen[3 * (long)k + e] = cen[3 * i + e];
    pdir[3 * (long)k + e] = d;
    if (palpha) palpha[3 * (long)k + e] = d * m;
  }
  pmag[k] = m;
  prv[k] = rv[i];
}

// contiguous results -> the pool slots of the live list
__global__ void k_wake_scatter(int n, const int *__restrict__ live, const double *__restrict__ pvel,
                               const double *__restrict__ pstr, double *vel, double *stretch) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  long i = live[k];
  for (int e = 0; e < 3; ++e) {
    vel[3 * i + e] = pvel[3 * (long)k + e];
    if (pstr) stretch[3 * i + e] = pstr[3 * (long)k + e];
  }
}

// ---- avoid_collision (mod_wake.f90:2350-2534) ------------------------------

// Body element seen by the collision avoidance: kind 0 surface panel, 1 vortex
// lattice (other kinds are skipped), centre, unit normal, largest edge and
// velocity.
struct PaElem {
  double cen[3], nor[3], ub[3];
  double maxedge;
  int kind, pad;
};

__global__ void k_pa_elems(int n, int kind, const int *__restrict__ nver,
                           const double *__restrict__ ver, const double *__restrict__ ub,
                           PaElem *out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  PaElem r{};
  r.kind = kind == DUST_ELEM_SURFPAN ? 0 : (kind == DUST_ELEM_RING ? 1 : -1);
  if (r.kind >= 0) {
    Panel g;
    panel_geo(nver[j] == 3 ? 3 : 4, ver + 12 * (long)j, g);
    for (int e = 0; e < 3; ++e) {
      r.cen[e] = g.cen[e];
      r.nor[e] = g.nor[e];
      r.ub[e] = ub ? ub[3 * (long)j + e] : 0.0;
    }
    r.maxedge = g.maxedge;
  }
  out[j] = r;
}

// The reference's expressions, each rounded as written (no fused
// multiply-add): the branch tests compare distances with thresholds.
__device__ inline double dot_rn(const double *a, const double *b) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ void pa_correct(int ne, const PaElem *__restrict__ E, const double pos1[3],
                           double r_vortex, const dust_pa_params &P, double vel[3]) {
#pragma clang fp contract(off)
  double blthick = r_vortex;
  const double r2d2 = sqrt(2.0) / 2.0;
  for (int ie = 0; ie < ne; ++ie) {
    const PaElem &el = E[ie];
    const int kind = el.kind;
    if (kind < 0) continue;
    double dist1[3], relvel[3], n[3];
    for (int k = 0; k < 3; ++k) {
      dist1[k] = pos1[k] - el.cen[k];
      n[k] = el.nor[k];
    }
    double elrad = el.maxedge * r2d2;
    double check_radius = P.dt * P.u_ref * P.rad_mult + elrad;
    double distn = sqrt(dot_rn(dist1, dist1));
    if (!(distn < check_radius)) continue;
    for (int k = 0; k < 3; ++k) relvel[k] = vel[k] - el.ub[k];
    double normvel = dot_rn(relvel, n);
    double dist1_nor, dt_part;
    if (kind == 0) {
      dist1_nor = dot_rn(dist1, n);
      if (dist1_nor < 0.0 || normvel >= 0.0) continue;
      dt_part = fmin(fmax(dist1_nor / (-normvel), 0.0), P.dt);
    } else {
      double q[3] = {dist1[0] * n[0], dist1[1] * n[1], dist1[2] * n[2]};
      dist1_nor = sqrt(dot_rn(q, q));
      if (dot_rn(dist1, n) * normvel >= 0.0) continue;
      dt_part = fmin(fmax(fabs(dist1_nor) / fabs(normvel), 0.0), P.dt);
    }
    double dist2[3];
    for (int k = 0; k < 3; ++k) dist2[k] = (pos1[k] + relvel[k] * dt_part) - el.cen[k];
    double dist2_nor;
    if (kind == 0) {
      dist2_nor = dot_rn(dist2, n);
    } else {
      double q[3] = {dist2[0] * n[0], dist2[1] * n[1], dist2[2] * n[2]};
      dist2_nor = sqrt(dot_rn(q, q));
      blthick = 2.0 * blthick;  // the reference doubles its local thickness here
    }
    if (!(dist2_nor <= blthick)) continue;
    double t2[3];
    for (int k = 0; k < 3; ++k) t2[k] = dist2[k] - n[k] * dist2_nor;
    if (!(sqrt(dot_rn(t2, t2)) < P.elrad_mult * elrad)) continue;
    double normvel_corr;
    if (dt_part < P.dt) {
      normvel_corr = -dist1_nor / P.dt_wake;
    } else {
      double dampf = dist2_nor / blthick;
      normvel_corr = normvel + (blthick - dist2_nor) / P.dt_wake * dampf;
    }
    double newvel[3];
    for (int k = 0; k < 3; ++k) newvel[k] = vel[k] + (normvel_corr - normvel) * n[k];
    normvel_corr = dot_rn(newvel, n);
    double nv = dot_rn(vel, n);
    double tanvel[3];
    for (int k = 0; k < 3; ++k) tanvel[k] = vel[k] - nv * n[k];
    double tanvel2 = dot_rn(tanvel, tanvel);
    double nveldiff = nv * nv - normvel_corr * normvel_corr + tanvel2;
    if (tanvel2 > 0.0 && normvel_corr * nv >= 0.0 && nveldiff >= 0.0) {
      double f = sqrt(nveldiff) / sqrt(tanvel2);
      for (int k = 0; k < 3; ++k) tanvel[k] = tanvel[k] * f;
    }
    for (int k = 0; k < 3; ++k) vel[k] = tanvel[k] + normvel_corr * n[k];
  }
}

// One thread per live slot; every thread walks the element list in order
// (uniform loads), almost always leaving at the check-radius test.
__global__ void __launch_bounds__(256)
    k_wake_avoid(int nmax, const unsigned char *__restrict__ freef,
                 const double *__restrict__ cen, const double *__restrict__ rv, double *vel,
                 int ne, const PaElem *__restrict__ E, dust_pa_params P) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmax || freef[i]) return;
  double pos[3] = {cen[3 * (long)i], cen[3 * (long)i + 1], cen[3 * (long)i + 2]};
  double v[3] = {vel[3 * (long)i], vel[3 * (long)i + 1], vel[3 * (long)i + 2]};
  pa_correct(ne, E, pos, rv[i], P, v);
  for (int k = 0; k < 3; ++k) vel[3 * (long)i + k] = v[k];
}

struct NotFree {
  __host__ __device__ bool operator()(const unsigned char &f) const { return f == 0; }
};
struct IsFree {
  __host__ __device__ bool operator()(const unsigned char &f) const { return f != 0; }
};

// New particle j (in order) inside the box takes free slot rank[j] of the
// ascending free list; rank = number of inside particles before j.
__global__ void k_wake_insert(int kmax, const int *__restrict__ kdev, const double *__restrict__ ncen,
                              const double *__restrict__ npv, const double *__restrict__ nrv,
                              const double *__restrict__ nvel, const int *__restrict__ free_list,
                              const int *__restrict__ n_free, dust_wake_params P,
                              unsigned char *freef, double *cen, double *dir, double *mag,
                              double *rv, double *vel, int *n_prt, int *status) {
  // a single block walks the (few) new particles in order
  __shared__ int s_rank;
  if (threadIdx.x == 0) s_rank = 0;
  __syncthreads();
  const int nf = *n_free;
  const int k = kdev ? min(*kdev, kmax) : kmax;
  for (int base = 0; base < k; base += blockDim.x) {
    int j = base + threadIdx.x;
    bool in = false;
    if (j < k) {
      in = true;
      for (int e = 0; e < 3; ++e) {
        double x = ncen[3 * j + e];
        in = in && x >= P.box_min[e] && x <= P.box_max[e];
      }
    }
    // exclusive rank of this particle among the inside ones of the chunk
    __shared__ int s_flag[1024];
    s_flag[threadIdx.x] = in ? 1 : 0;
    __syncthreads();
    int r = s_rank;
    for (int t = 0; t < (int)threadIdx.x; ++t) r += s_flag[t];
    __syncthreads();
    if (in) {
      if (r < nf) {
        int slot = free_list[r];
        double pv[3] = {npv[3 * j], npv[3 * j + 1], npv[3 * j + 2]};
        double m = sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]);
        freef[slot] = 0;
        mag[slot] = m;
        for (int e = 0; e < 3; ++e) {
          dir[3 * (long)slot + e] = m > 1.0e-13 ? pv[e] / m : pv[e];
          cen[3 * (long)slot + e] = ncen[3 * j + e];
          vel[3 * (long)slot + e] = nvel[3 * j + e];
        }
        rv[slot] = nrv[j];
        atomicAdd(n_prt, 1);
      } else {
        atomicOr(status, 1);  // more new particles than free slots
      }
    }
    if (threadIdx.x == blockDim.x - 1) s_rank = r + (in ? 1 : 0);
    __syncthreads();
  }
}

struct WakeWs {
  DevBuf<PaElem> pa;
  DevBuf<int> list, count;
  DevBuf<char> tmp;
  HostBuf<int> hcount;
};

WakeWs g_wws;

}  // namespace
void wake_release() {
  g_wws.pa.release();
  g_wws.list.release();
  g_wws.count.release();
  g_wws.tmp.release();
  g_wws.hcount.release();
}
namespace {

int select_slots(int nmax, const unsigned char *freef, bool want_free, int *out, int *d_count,
                 hipStream_t st) {
  hipcub::CountingInputIterator<int> it(0);
  size_t bytes = 0;
  if (want_free) {
    hipcub::TransformInputIterator<bool, IsFree, const unsigned char *> fl(freef, IsFree());
    DUST_HIP(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, fl, out, d_count, nmax, st));
    DUST_TRY(g_wws.tmp.reserve(bytes + 256));
    DUST_HIP(hipcub::DeviceSelect::Flagged(g_wws.tmp.p, bytes, it, fl, out, d_count, nmax, st));
  } else {
    hipcub::TransformInputIterator<bool, NotFree, const unsigned char *> fl(freef, NotFree());
    DUST_HIP(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, fl, out, d_count, nmax, st));
    DUST_TRY(g_wws.tmp.reserve(bytes + 256));
    DUST_HIP(hipcub::DeviceSelect::Flagged(g_wws.tmp.p, bytes, it, fl, out, d_count, nmax, st));
  }
  return 0;
}

}  // namespace
}  // namespace dust

namespace dust {
// Insertion of the first min(*kdev, kmax) new particles (all kmax when kdev
// is NULL) into the lowest free slots, in order; `status` (device int) is
// or-ed with 1 when the free slots run out.  No host synchronisation.
int wake_insert_dev(int nmax, unsigned char *free_flag, double *cen, double *dir, double *mag,
                    double *r_vortex, double *vel, int kmax, const int *kdev,
                    const double *new_cen, const double *new_partvec, const double *new_rv,
                    const double *new_vel, const dust_wake_params *prm, int *n_prt, int *status,
                    hipStream_t st) {
  if (kmax <= 0) return 0;
  DUST_TRY(g_wws.list.reserve(nmax));
  DUST_TRY(g_wws.count.reserve(2));
  DUST_TRY(select_slots(nmax, free_flag, true, g_wws.list.p, g_wws.count.p, st));
  k_wake_insert<<<1, 1024, 0, st>>>(kmax, kdev, new_cen, new_partvec, new_rv, new_vel,
                                   g_wws.list.p, g_wws.count.p, *prm, free_flag, cen, dir, mag,
                                   r_vortex, vel, n_prt, status);
  DUST_HIP(hipGetLastError());
  return 0;
}
}  // namespace dust

using namespace dust;

extern "C" {

int dust_wake_update_dev(int nmax, unsigned char *free_flag, double *cen, double *dir,
                         const double *mag, const double *vel, double *stretch,
                         const double *rotu, const dust_wake_params *prm, int *n_prt,
                         void *stream) {
  DUST_CHECK(prm && nmax >= 0 && n_prt, DUST_E_ARG, "invalid wake update arguments");
  if (nmax == 0) return 0;
  DUST_CHECK(free_flag && cen && dir && mag && vel, DUST_E_ARG, "missing particle arrays");
  DUST_CHECK(!(prm->use_vs || prm->use_vd) || stretch, DUST_E_ARG, "stretch needed");
  DUST_CHECK(!((prm->use_vs || prm->use_vd) && prm->use_divfilt) || rotu, DUST_E_ARG,
             "rotu needed for the divergence filter");
  hipStream_t st = (hipStream_t)stream;
  k_wake_update<<<ceil_div(nmax, 256), 256, 0, st>>>(nmax, free_flag, cen, dir, mag, vel, stretch,
                                                     rotu, *prm, n_prt);
  DUST_HIP(hipGetLastError());
  return 0;
}

int dust_wake_live_dev(int nmax, const unsigned char *free_flag, int *live, int *n_live,
                       void *stream) {
  DUST_CHECK(nmax >= 0 && n_live && (nmax == 0 || (free_flag && live)), DUST_E_ARG,
             "invalid live-list arguments");
  hipStream_t st = (hipStream_t)stream;
  if (nmax == 0) {
    DUST_HIP(hipMemsetAsync(n_live, 0, sizeof(int), st));
    return 0;
  }
  return select_slots(nmax, free_flag, false, live, n_live, st);
}

int dust_wake_shed_dev(int nmax, unsigned char *free_flag, double *cen, double *dir,
                       double *mag, double *r_vortex, double *vel, int k, const double *new_cen,
                       const double *new_partvec, const double *new_rv, const double *new_vel,
                       const dust_wake_params *prm, int *n_prt, void *stream) {
  DUST_CHECK(prm && nmax >= 0 && k >= 0 && n_prt, DUST_E_ARG, "invalid shedding arguments");
  if (k == 0) return 0;
  DUST_CHECK(new_cen && new_partvec && new_rv && new_vel, DUST_E_ARG, "missing new particles");
  DUST_CHECK(nmax > 0 && free_flag && cen && dir && mag && r_vortex && vel, DUST_E_ARG,
             "missing particle arrays");
  hipStream_t st = (hipStream_t)stream;
  DUST_TRY(g_wws.list.reserve(nmax));
  DUST_TRY(g_wws.count.reserve(2));
  DUST_TRY(g_wws.hcount.reserve(1));
  DUST_TRY(select_slots(nmax, free_flag, true, g_wws.list.p, g_wws.count.p, st));
  DUST_HIP(hipMemsetAsync(g_wws.count.p + 1, 0, sizeof(int), st));
  k_wake_insert<<<1, 1024, 0, st>>>(k, nullptr, new_cen, new_partvec, new_rv, new_vel, g_wws.list.p,
                                   g_wws.count.p, *prm, free_flag, cen, dir, mag, r_vortex, vel,
                                   n_prt, g_wws.count.p + 1);
  DUST_HIP(hipGetLastError());
  DUST_HIP(hipMemcpyAsync(g_wws.hcount.p, g_wws.count.p + 1, sizeof(int), hipMemcpyDeviceToHost,
                          st));
  DUST_HIP(hipStreamSynchronize(st));
  DUST_CHECK(g_wws.hcount.p[0] == 0, DUST_E_CAPACITY,
             "Exceeding the maximum number of %d wake particles introduced", nmax);
  return 0;
}

int dust_wake_gather_dev(int n, const int *live, const double *cen, const double *dir,
                         const double *mag, const double *r_vortex, double *pcen, double *pdir,
                         double *pmag, double *prv, double *palpha, void *stream) {
  DUST_CHECK(n >= 0, DUST_E_ARG, "negative count");
  if (n == 0) return 0;
  DUST_CHECK(live && cen && dir && mag && r_vortex && pcen && pdir && pmag && prv, DUST_E_ARG,
             "missing arrays");
  k_wake_gather<<<ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(n, live, cen, dir, mag,
                                                                   r_vortex, pcen, pdir, pmag,
                                                                   prv, palpha);
  DUST_HIP(hipGetLastError());
  return 0;
}

int dust_wake_scatter_dev(int n, const int *live, const double *pvel, const double *pstretch,
                          double *vel, double *stretch, void *stream) {
  DUST_CHECK(n >= 0, DUST_E_ARG, "negative count");
  if (n == 0) return 0;
  DUST_CHECK(live && pvel && vel && (!pstretch || stretch), DUST_E_ARG, "missing arrays");
  k_wake_scatter<<<ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(n, live, pvel, pstretch, vel,
                                                                    stretch);
  DUST_HIP(hipGetLastError());
  return 0;
}

int dust_wake_avoid_collision_dev(int nmax, const unsigned char *free_flag, const double *cen,
                                  const double *r_vortex, double *vel, int ngroups,
                                  const dust_elem_group *groups, const double *const *ub,
                                  const dust_pa_params *prm, void *stream) {
  DUST_CHECK(prm && nmax >= 0 && ngroups >= 0 && (ngroups == 0 || groups), DUST_E_ARG,
             "invalid collision-avoidance arguments");
  long ne = 0;
  for (int k = 0; k < ngroups; ++k) {
    DUST_CHECK(groups[k].n >= 0 && (groups[k].n == 0 || groups[k].kind == DUST_ELEM_LINE ||
                                    (groups[k].ver && groups[k].n_ver)),
               DUST_E_ARG, "element group %d: missing vertices", k);
    ne += groups[k].n;
  }
  if (nmax == 0 || ne == 0) return 0;
  DUST_CHECK(free_flag && cen && r_vortex && vel, DUST_E_ARG, "missing particle arrays");
  hipStream_t st = (hipStream_t)stream;
  DUST_TRY(g_wws.pa.reserve(ne));
  long off = 0;
  for (int k = 0; k < ngroups; ++k) {
    const dust_elem_group &G = groups[k];
    if (G.n == 0) continue;
    k_pa_elems<<<ceil_div(G.n, 128), 128, 0, st>>>(G.n, G.kind, G.n_ver, G.ver,
                                                   ub ? ub[k] : nullptr, g_wws.pa.p + off);
    off += G.n;
  }
  k_wake_avoid<<<ceil_div(nmax, 256), 256, 0, st>>>(nmax, free_flag, cen, r_vortex, vel, (int)ne,
                                                    g_wws.pa.p, *prm);
  DUST_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
