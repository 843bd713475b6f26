// Octree FMM for vortex particles on gfx950: tree build, expansions and the
// plain FP64 evaluation, plus the C ABI of dust_fmm_*.
//
// Reference behaviour: src/fmm/mod_octree.f90 (initialize_octree 251-518,
// sort_particles 703-996, calculate_multipole 1002-1133, apply_multipole
// 1139-1429) and src/fmm/mod_multipole.f90 (P2M 224, M2M 256, M2L 282,
// L2L 360, compute_der_ker 388, set_degree_polyexp 454).  Particle kernels:
// src/aeroel/mod_vortpart.f90:139-336.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "fmm_eval.h"

namespace dust {

// ---------------------------------------------------------------------------
// Error text
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}


static void mono_powers(int deg, std::vector<int> &pwr) {
  pwr.clear();
  for (int o = 0; o <= deg; ++o)
    for (int i = o; i >= 0; --i)
      for (int j = o - i; j >= 0; --j) {
        pwr.push_back(i);
        pwr.push_back(j);
        pwr.push_back(o - i - j);
      }
}

// Taylor coefficients of the Rosenhead-regularised Green's function
// 1/(4 pi sqrt(|x|^2 + delta^2)) at x = diff, by the Lindsay-Krasny recurrence
// (mod_multipole.f90:409-434), and the derived velocity (D) and gradient (Dc)
// tables (436-449).  Writes the first nU entries of D and Dc in the SoA order
// used on the device: out[e*nU + m] for e = 0..2 (D) and e = 3 + c + 3 i (Dc).
static void kernel_derivatives(int p, const double diff[3], double delta, int nU, double *out) {
  int deg = 2 * p + 2;
  int nb = mono_count(deg);
  std::vector<double> bk(nb, 0.0);
  double R2 = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2] + delta * delta;
  for (int k = 0; k <= deg; ++k)
    for (int j = 0; j <= deg - k; ++j)
      for (int i = 0; i <= deg - j - k; ++i) {
        if (i == 0 && j == 0 && k == 0) {
          bk[0] = 1.0 / (4.0 * kPi * std::sqrt(R2));
          continue;
        }
        double km = (double)(i + j + k), s1 = 0.0, s2 = 0.0;
        if (i >= 1) s1 += diff[0] * bk[mono_index(i - 1, j, k)];
        if (j >= 1) s1 += diff[1] * bk[mono_index(i, j - 1, k)];
        if (k >= 1) s1 += diff[2] * bk[mono_index(i, j, k - 1)];
        if (i >= 2) s2 += bk[mono_index(i - 2, j, k)];
        if (j >= 2) s2 += bk[mono_index(i, j - 2, k)];
        if (k >= 2) s2 += bk[mono_index(i, j, k - 2)];
        bk[mono_index(i, j, k)] = 1.0 / (R2 * km) * ((2.0 * km - 1.0) * s1 - (km - 1.0) * s2);
      }
  int nD = mono_count(deg - 1);
  std::vector<double> D(3 * nD, 0.0);
  for (int k = 0; k <= deg - 1; ++k)
    for (int j = 0; j <= deg - 1 - k; ++j)
      for (int i = 0; i <= deg - 1 - j - k; ++i) {
        int m = mono_index(i, j, k);
        D[3 * m + 0] = -(1.0 + i) * bk[mono_index(i + 1, j, k)];
        D[3 * m + 1] = -(1.0 + j) * bk[mono_index(i, j + 1, k)];
        D[3 * m + 2] = -(1.0 + k) * bk[mono_index(i, j, k + 1)];
      }
  for (int m = 0; m < nU; ++m)
    for (int c = 0; c < 3; ++c) out[c * nU + m] = D[3 * m + c];
  // Dc(:,ii,m) = -(1+pwr_ii(m)) D(:, m + e_ii); only orders <= 2p are needed.
  std::vector<int> pw;
  mono_powers(2 * p, pw);
  for (int m = 0; m < nU; ++m) {
    int i = pw[3 * m], j = pw[3 * m + 1], k = pw[3 * m + 2];
    int mx = mono_index(i + 1, j, k), my = mono_index(i, j + 1, k), mz = mono_index(i, j, k + 1);
    for (int c = 0; c < 3; ++c) {
      out[(3 + c + 0) * nU + m] = -(1.0 + i) * D[3 * mx + c];
      out[(3 + c + 3) * nU + m] = -(1.0 + j) * D[3 * my + c];
      out[(3 + c + 6) * nU + m] = -(1.0 + k) * D[3 * mz + c];
    }
  }
}

// ---------------------------------------------------------------------------
// Device helpers for the cell numbering
TreeGeom make_geom(const FmmHandle *h) {
  TreeGeom G;
  memset(&G, 0, sizeof(G));
  G.nlev = h->nlev;
  for (int d = 0; d < 3; ++d) {
    G.nbox[d] = h->prm.nbox[d];
    G.xmin[d] = h->xmin[d];
    G.xmax[d] = h->xmax[d];
  }
  for (int l = 0; l < h->nlev; ++l) {
    for (int d = 0; d < 3; ++d) {
      G.ncl[l][d] = h->lev[l].ncl[d];
      G.kdl[l][d] = h->lev[l].kdl[d];
    }
    G.off[l] = h->lev[l].off;
    G.kd_base[l] = h->lev[l].kd_base;
  }
  G.csize_fine = h->lev[h->nlev - 1].csize;
  if (h->gr.active) {
    G.dc = h->gr.dc.p;
    G.act_cap = h->gr.act_cap;
    G.leaf_cap = h->gr.leaf_cap;
    G.item_cap = h->gr.item_cap;
  }
  return G;
}

// ---------------------------------------------------------------------------
// Tree build kernels (sort_particles, mod_octree.f90:703-996)
// ---------------------------------------------------------------------------

// Finest-level cell of every particle: idx = ceiling((x - xmin)/csize)
// (mod_octree.f90:755), with the out-of-box check of 758-762.
__global__ void k_keys(TreeGeom G, int n, const double *__restrict__ cen, unsigned *key, int *idx,
                       int *err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (G.dc) n = G.dc[kGcN];
  if (i >= n) return;
  int L = G.nlev - 1;
  int id[3];
  bool bad = false;
  for (int d = 0; d < 3; ++d) {
    double q = ceil((cen[3 * i + d] - G.xmin[d]) / G.csize_fine);
    // compare in floating point first so huge values cannot overflow the int
    if (!(q >= 1.0 && q <= (double)G.ncl[L][d])) bad = true;
    id[d] = bad ? 1 : (int)q;
  }
  if (bad) {
    atomicMin(err, i);
    key[i] = 0;
  } else {
    key[i] = (unsigned)cell_encode(G, L, id[0] - 1, id[1] - 1, id[2] - 1);
  }
  idx[i] = i;
}

// Gather the particle data into sorted SoA arrays.
__global__ void k_gather(int n, const int *__restrict__ perm, const double *__restrict__ cen,
                         const double *__restrict__ dir, const double *__restrict__ mag,
                         const double *__restrict__ rv, double *px, double *py, double *pz,
                         double *dx, double *dy, double *dz, double *pm, double *pr,
                         const int *__restrict__ n_dev) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (n_dev) n = *n_dev;
  if (i >= n) return;
  int q = perm[i];
  px[i] = cen[3 * q];
  py[i] = cen[3 * q + 1];
  pz[i] = cen[3 * q + 2];
  dx[i] = dir[3 * q];
  dy[i] = dir[3 * q + 1];
  dz[i] = dir[3 * q + 2];
  pm[i] = mag[q];
  pr[i] = rv[q];
}

// fstart[c] = first sorted particle with key >= c, for c = 0..nfine (the
// particle count read from n_dev when given: after a redistribution).
__global__ void k_fstart(long nfine, int n, const int *__restrict__ n_dev,
                         const unsigned *__restrict__ ks, int *fstart) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > nfine) return;
  if (n_dev) n = *n_dev;
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if ((long)ks[mid] < c) lo = mid + 1; else hi = mid;
  }
  fstart[c] = lo;
}

// ---------------------------------------------------------------------------
// Particle redistribution (sort_particles, mod_octree.f90:827-894)
// ---------------------------------------------------------------------------

// Element centres binned with the particles' ceiling rule at the finest level
// (mod_octree.f90:777-800; centres outside the box are skipped, the warning
// branch of 784-795) and marked on their ancestor at level ls (the upward
// panel counts of 801-817, of which only "> 0" is read).
__global__ void k_pr_bin(TreeGeom G, int m, const double *__restrict__ cen, int ls,
                         unsigned char *has) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  int L = G.nlev - 1;
  int id[3];
  for (int d = 0; d < 3; ++d) {
    double q = ceil((cen[3 * e + d] - G.xmin[d]) / G.csize_fine);
    if (!(q >= 1.0 && q <= (double)G.ncl[L][d])) return;
    id[d] = (int)q - 1;
  }
  has[cell_encode(G, L, id[0], id[1], id[2]) >> (3 * (L - ls))] = 1;
}

// near_wall of the level-ls cells: an element in the cell or in one of its
// neighbours inside the grid (mod_octree.f90:829-843).
__global__ void k_pr_near(TreeGeom G, int ls, long nc, const unsigned char *__restrict__ has,
                          unsigned char *near) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int ix, iy, iz;
  cell_decode(G, ls, c, ix, iy, iz);
  unsigned char v = 0;
  for (int kc = -1; kc <= 1; ++kc)
    for (int jc = -1; jc <= 1; ++jc)
      for (int ic = -1; ic <= 1; ++ic) {
        int a = ix + ic, b = iy + jc, d = iz + kc;
        if (a < 0 || b < 0 || d < 0 || a >= G.ncl[ls][0] || b >= G.ncl[ls][1] ||
            d >= G.ncl[ls][2])
          continue;
        v |= has[cell_encode(G, ls, a, b, d)];
      }
  near[c] = v;
}

// The redistribution loop of one finest cell (mod_octree.f90:845-891), one
// thread per cell over its sorted particles, which keep the push order of the
// reference's lists.  Walking the list, a particle with
// mag < (1/ratio) * (sum of the cell's mag) / npart is freed; redistr =
// (mag * dir) / nprt, nprt counting the freed particle, is added to each
// other particle still in the cell, whose mag = norm2(vort) and
// dir = vort / norm2(vort) (0 for a zero vector) are recomputed.  Pulling the
// freed particle from the list leaves the next one at the same position, so
// one ordered pass over the positions visits the particles as the reference's
// while-loop does.  Unfused FP64 arithmetic, norm2 = sqrt(x*x + y*y + z*z).
__global__ void k_pr_cells(TreeGeom G, int ls, const int *__restrict__ fstart,
                           const unsigned char *__restrict__ near, double inv_ratio, double *pm,
                           double *dx, double *dy, double *dz, unsigned char *keep) {
#pragma clang fp contract(off)
  int L = G.nlev - 1;
  long nc = (long)G.ncl[L][0] * G.ncl[L][1] * G.ncl[L][2];
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int s = fstart[c], e = fstart[c + 1];
  for (int t = s; t < e; ++t) keep[t] = 1;
  if (e == s || near[c >> (3 * (L - ls))]) return;
  double ave = 0.0;
  for (int t = s; t < e; ++t) ave += pm[t];
  ave = ave / (double)(e - s);
  const double thr = inv_ratio * ave;
  int cnt = e - s;
  for (int t = s; t < e; ++t) {
    const double mt = pm[t];
    if (!(mt < thr)) continue;
    const double rx = (mt * dx[t]) / (double)cnt;
    const double ry = (mt * dy[t]) / (double)cnt;
    const double rz = (mt * dz[t]) / (double)cnt;
    keep[t] = 0;
    --cnt;
    for (int q = s; q < e; ++q) {
      if (!keep[q]) continue;
      const double mq = pm[q];
      const double vx = mq * dx[q] + rx, vy = mq * dy[q] + ry, vz = mq * dz[q] + rz;
      const double vn = sqrt(vx * vx + vy * vy + vz * vz);
      pm[q] = vn;
      dx[q] = vn > 0.0 ? vx / vn : 0.0;
      dy[q] = vn > 0.0 ? vy / vn : 0.0;
      dz[q] = vn > 0.0 ? vz / vn : 0.0;
    }
  }
}

// Outcome of the redistribution in input order: kept flag and strengths.
__global__ void k_pr_out(int n, const int *__restrict__ perm, const unsigned char *__restrict__ ks,
                         const double *__restrict__ pm, const double *__restrict__ dx,
                         const double *__restrict__ dy, const double *__restrict__ dz,
                         unsigned char *keep, double *dir, double *mag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int q = perm[i];
  keep[q] = ks[i];
  mag[q] = pm[i];
  dir[3 * q] = dx[i];
  dir[3 * q + 1] = dy[i];
  dir[3 * q + 2] = dz[i];
}

// The kept particles moved to the front of the sorted arrays, in order
// (sel = their sorted positions, *ns of them).
struct PrArrays {
  const double *src[8];
  double *dst[8];
};
__global__ void k_pr_compact(int n, const int *__restrict__ ns, const int *__restrict__ sel,
                             PrArrays A, const unsigned *__restrict__ ks, unsigned *ks_out,
                             const int *__restrict__ perm, int *perm_out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || j >= *ns) return;
  int i = sel[j];
  for (int a = 0; a < 8; ++a) A.dst[a][j] = A.src[a][i];
  ks_out[j] = ks[i];
  perm_out[j] = perm[i];
}

// Outputs of the freed particles (no evaluation) set to zero.
__global__ void k_pr_zero(int n, const unsigned char *__restrict__ keep, double *v, double *s,
                          double *tv, double *rot) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n || keep[q]) return;
  v[3 * q] = v[3 * q + 1] = v[3 * q + 2] = 0.0;
  s[3 * q] = s[3 * q + 1] = s[3 * q + 2] = 0.0;
  tv[q] = 0.0;
  if (rot) rot[3 * q] = rot[3 * q + 1] = rot[3 * q + 2] = 0.0;
}

// Finest level: leaf iff npart >= min_part, or the tree has one level
// (mod_octree.f90:898-907).
__global__ void k_flags_fine(TreeGeom G, const int *__restrict__ fstart, int min_part,
                             unsigned char *flags) {
  int L = G.nlev - 1;
  long nc = (long)G.ncl[L][0] * G.ncl[L][1] * G.ncl[L][2];
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int np = fstart[c + 1] - fstart[c];
  flags[G.off[L] + c] = (np >= min_part || G.nlev == 1) ? kLeaf : 0;
}

// One level upwards (mod_octree.f90:911-961): a cell with a leaf or branch
// child becomes a branch and its other children become leaves; otherwise its
// children are de-activated and it is a leaf if it holds min_part particles
// (always on level 1).
__global__ void k_flags_level(TreeGeom G, int lvl, const int *__restrict__ fstart, int min_part,
                              unsigned char *flags) {
  long nc = (long)G.ncl[lvl][0] * G.ncl[lvl][1] * G.ncl[lvl][2];
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  unsigned char *ch = &flags[G.off[lvl + 1] + 8 * c];
  bool got = false;
  for (int p = 0; p < 8; ++p) got |= (ch[p] != 0);
  if (got) {
    for (int p = 0; p < 8; ++p)
      if (ch[p] == 0) ch[p] = kLeaf;
    flags[G.off[lvl] + c] = kBranch;
  } else {
    for (int p = 0; p < 8; ++p) ch[p] = 0;
    int2 r = cell_range(G, fstart, lvl, c);
    flags[G.off[lvl] + c] = (r.y - r.x >= min_part || lvl == 0) ? kLeaf : 0;
  }
}

__global__ void k_fill_int(long n, int *p, int v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u8(long n, unsigned char *p, unsigned char v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_slots(int n_active, const int *__restrict__ active, int *slot,
                        const int *__restrict__ dc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (dc) n_active = dc[kGcOvf] ? 0 : dc[kGcActive];
  if (i < n_active) slot[active[i]] = i;
}

// First position of each level in the ascending active list.
__global__ void k_level_starts(TreeGeom G, int n_active, const int *__restrict__ active, int *out) {
  int l = threadIdx.x;
  if (l > G.nlev) return;
  if (G.dc) n_active = min(G.dc[kGcActive], G.act_cap);
  long key = (l == G.nlev) ? 0x7fffffffffffL : G.off[l];
  int lo = 0, hi = n_active;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if ((long)active[mid] < key) lo = mid + 1; else hi = mid;
  }
  out[l] = lo;
}

// P2P source list of every leaf, in the order of apply_multipole
// (mod_octree.f90:1258-1348): the 26 same-level neighbours that are active
// (their whole subtree), then leaf_neigh = leaves among the neighbours of
// every ancestor from level 1 down to the parent (964-988), then the leaf
// itself (the target is skipped by index).  Each entry is a range of sorted
// particles.
__global__ void __launch_bounds__(256)
    k_p2p_lists(TreeGeom G, int n_leaves, const int *__restrict__ leaves,
                const unsigned char *__restrict__ flags, const int *__restrict__ fstart, int cap,
                int2 *rng, int *cnt, int *tot) {
  // one wavefront per leaf: lane q tests candidate q of the list in order
  // (the 26 same-level offsets in (k, j, i) order, then 26 per ancestor
  // level); the hits are compacted in that order by ballot
  const int lane = threadIdx.x & 63;
  const int lv = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (G.dc) n_leaves = g_stop(G) ? 0 : G.dc[kGcLeaves];
  if (lv >= n_leaves) return;
  long g = leaves[lv];
  int lvl = lvl_of(G, g);
  long c = g - G.off[lvl];
  int ix, iy, iz;
  cell_decode(G, lvl, c, ix, iy, iz);
  int2 *out = &rng[(long)lv * cap];
  int w = 0;
  long T = 0;
  const int ncand = 26 * (1 + lvl);
  for (int base = 0; base < ncand; base += 64) {
    const int q = base + lane;
    int2 r = make_int2(0, 0);
    bool ok = false;
    if (q < ncand) {
      const int blk = q / 26, idx = q - 26 * blk;
      const int t = idx < 13 ? idx : idx + 1;  // skip the centre offset
      const int ii = t % 3 - 1, jj = (t / 3) % 3 - 1, kk = t / 9 - 1;
      // block 0: the leaf's own level; block 1 + al: ancestor level al
      const int al = blk == 0 ? lvl : blk - 1;
      const int sh = lvl - al;
      const int a = (ix >> sh) + ii, b = (iy >> sh) + jj, d = (iz >> sh) + kk;
      if (a >= 0 && b >= 0 && d >= 0 && a < G.ncl[al][0] && b < G.ncl[al][1] &&
          d < G.ncl[al][2]) {
        const long nc = cell_encode(G, al, a, b, d);
        const unsigned char f = flags[G.off[al] + nc];
        if (blk == 0 ? f != 0 : f == kLeaf) {
          r = cell_range(G, fstart, al, nc);
          ok = r.y > r.x;
        }
      }
    }
    const unsigned long long m = __ballot(ok);
    if (ok) {
      const int pos = w + __popcll(m & ((1ull << lane) - 1ull));
      out[pos] = r;
      T += r.y - r.x;
    }
    w += __popcll(m);
  }
  for (int o = 32; o > 0; o >>= 1) T += __shfl_xor(T, o, 64);
  if (lane == 0) {
    const int2 own = cell_range(G, fstart, lvl, c);
    out[w] = own;
    cnt[lv] = w + 1;
    tot[lv] = (int)(T + (own.y - own.x));
  }
}

// ---------------------------------------------------------------------------
// Expansions (calculate_multipole, mod_octree.f90:1002-1133)
// ---------------------------------------------------------------------------
__device__ inline double ipowd(double x, int e) {
  double r = 1.0;
  for (int q = 0; q < e; ++q) r *= x;
  return r;
}

// P2M on every leaf (leaf_M_multipole, mod_multipole.f90:224-251):
// a(:,m) = sum_i mag_i (x_i - c)^pwr(m) dir_i, particles in list order.
// One block per leaf, one thread per monomial; particle powers staged in LDS.
constexpr int kP2MChunk = 64;
__global__ void k_p2m(TreeGeom G, int deg, int nmon, const int *__restrict__ pwr,
                      const int *__restrict__ leaves, const int *__restrict__ slot,
                      const int *__restrict__ fstart, const double *__restrict__ px,
                      const double *__restrict__ py, const double *__restrict__ pz,
                      const double *__restrict__ dx, const double *__restrict__ dy,
                      const double *__restrict__ dz, const double *__restrict__ pm, double *mp_a) {
  extern __shared__ double sm[];
  if (G.dc && (g_stop(G) || (int)blockIdx.x >= G.dc[kGcLeaves])) return;
  const int P1 = deg + 1;
  double *pw = sm;                         // [kP2MChunk][3][P1]
  double *al = sm + kP2MChunk * 3 * P1;    // [kP2MChunk][4] mag, dir
  long g = leaves[blockIdx.x];
  int lvl = lvl_of(G, g);
  long c = g - G.off[lvl];
  int ix, iy, iz;
  cell_decode(G, lvl, c, ix, iy, iz);
  double cx = cell_center(G, lvl, 0, ix), cy = cell_center(G, lvl, 1, iy),
         cz = cell_center(G, lvl, 2, iz);
  int2 r = cell_range(G, fstart, lvl, c);
  int m = threadIdx.x;
  int e0 = 0, e1 = 0, e2 = 0;
  if (m < nmon) { e0 = pwr[3 * m]; e1 = pwr[3 * m + 1]; e2 = pwr[3 * m + 2]; }
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int base = r.x; base < r.y; base += kP2MChunk) {
    int cnt = min(kP2MChunk, r.y - base);
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      int q = base + t;
      double d0 = px[q] - cx, d1 = py[q] - cy, d2 = pz[q] - cz;
      double v0 = 1.0, v1 = 1.0, v2 = 1.0;
      for (int e = 0; e < P1; ++e) {
        pw[(t * 3 + 0) * P1 + e] = v0;
        pw[(t * 3 + 1) * P1 + e] = v1;
        pw[(t * 3 + 2) * P1 + e] = v2;
        v0 *= d0; v1 *= d1; v2 *= d2;
      }
      al[4 * t] = pm[q];
      al[4 * t + 1] = dx[q];
      al[4 * t + 2] = dy[q];
      al[4 * t + 3] = dz[q];
    }
    __syncthreads();
    if (m < nmon) {
      for (int t = 0; t < cnt; ++t) {
        double pr = pw[(t * 3 + 0) * P1 + e0] * pw[(t * 3 + 1) * P1 + e1] * pw[(t * 3 + 2) * P1 + e2];
        double f = al[4 * t] * pr;
        a0 += f * al[4 * t + 1];
        a1 += f * al[4 * t + 2];
        a2 += f * al[4 * t + 3];
      }
    }
    __syncthreads();
  }
  if (m < nmon) {
    double *a = &mp_a[(long)slot[g] * 3 * nmon];
    a[3 * m] = a0; a[3 * m + 1] = a1; a[3 * m + 2] = a2;
  }
}

// M2M into the branches of one level (M2M_multipole, mod_multipole.f90:256-278):
// a(:,m) = sum_children sum_{s<=m} C(m,s) (c_child - c)^(m-s) a_child(:,s).
__global__ void k_m2m(TreeGeom G, int lvl, int deg, int nmon, const int *__restrict__ pwr,
                      const double *__restrict__ fact, const int *__restrict__ active, int a0i,
                      const unsigned char *__restrict__ flags, const int *__restrict__ slot,
                      double *mp_a) {
  __shared__ double dp[8][3][11];
  __shared__ double sB[11 * 11];  // binomials fact[n] / (fact[r] fact[n-r]), as below
  if (G.dc) {
    if (g_stop(G)) return;
    a0i = G.dc[kGcActStart + lvl];
    if ((int)blockIdx.x >= G.dc[kGcActStart + lvl + 1] - a0i) return;
  }
  long g = active[a0i + blockIdx.x];
  if (flags[g] != kBranch) return;
  for (int t = threadIdx.x; t < (deg + 1) * (deg + 1); t += blockDim.x) {
    const int n = t / (deg + 1), r = t - n * (deg + 1);
    sB[t] = r <= n ? fact[n] / (fact[r] * fact[n - r]) : 0.0;
  }
  long c = g - G.off[lvl];
  int ix, iy, iz;
  cell_decode(G, lvl, c, ix, iy, iz);
  double cc[3] = {cell_center(G, lvl, 0, ix), cell_center(G, lvl, 1, iy),
                  cell_center(G, lvl, 2, iz)};
  for (int t = threadIdx.x; t < 8; t += blockDim.x) {
    int p = t;
    int ci[3] = {2 * ix + (p & 1), 2 * iy + ((p >> 1) & 1), 2 * iz + ((p >> 2) & 1)};
    for (int d = 0; d < 3; ++d) {
      double dd = cell_center(G, lvl + 1, d, ci[d]) - cc[d];
      double v = 1.0;
      for (int e = 0; e <= deg; ++e) { dp[p][d][e] = v; v *= dd; }
    }
  }
  // the children's expansions (8 x 3 nmon doubles), loaded together
  extern __shared__ double sA[];
  for (int q = threadIdx.x; q < 8 * 3 * nmon; q += blockDim.x) {
    const int p = q / (3 * nmon), r = q - p * 3 * nmon;
    sA[q] = mp_a[(long)slot[G.off[lvl + 1] + 8 * c + p] * 3 * nmon + r];
  }
  __syncthreads();
  int m = threadIdx.x;
  if (m >= nmon) return;
  int i = pwr[3 * m], j = pwr[3 * m + 1], k = pwr[3 * m + 2];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int p = 0; p < 8; ++p) {
    const double *ac = &sA[p * 3 * nmon];
    for (int ks = 0; ks <= k; ++ks)
      for (int js = 0; js <= j; ++js)
        for (int is = 0; is <= i; ++is) {
          int s = mono_index(is, js, ks);
          const int w = deg + 1;
          double nb = sB[i * w + is] * sB[j * w + js] * sB[k * w + ks];
          double f = nb * (dp[p][0][i - is] * dp[p][1][j - js] * dp[p][2][k - ks]);
          s0 += f * ac[3 * s];
          s1 += f * ac[3 * s + 1];
          s2 += f * ac[3 * s + 2];
        }
  }
  double *a = &mp_a[(long)slot[g] * 3 * nmon];
  a[3 * m] = s0; a[3 * m + 1] = s1; a[3 * m + 2] = s2;
}

// Offset of the kernel-derivative table for displacement (di,dj,dk) on level lvl.
__device__ inline long kd_table(const TreeGeom &G, int lvl, int di, int dj, int dk) {
  int w0 = 2 * G.kdl[lvl][0] + 1, w1 = 2 * G.kdl[lvl][1] + 1;
  return G.kd_base[lvl] + (di + G.kdl[lvl][0]) + w0 * ((dj + G.kdl[lvl][1]) + (long)w1 * (dk + G.kdl[lvl][2]));
}

// M2L on the active cells of one level (M2L_multipole, mod_multipole.f90:282-356),
// plain FP64.  Interaction list in the order of initialize_octree
// (mod_octree.f90:422-476).  One block per target, one thread per output
// monomial m; the derivative table of each source's offset and the source
// expansion are staged in LDS.
//   b(:,m) += (-1)^|m| sum_n mult(m,n) D(:,m+n) x a(:,n)
//   c(i,:,m) += (-1)^|m| sum_n mult(m,n) Dc(:,i,m+n) x a(:,n)
__global__ void k_m2l_fp64(TreeGeom G, int lvl, int nmon, int nU, int use_vs,
                           const int *__restrict__ pwr, const double *__restrict__ multT,
                           const int *__restrict__ mnT, const double *__restrict__ kd,
                           const int *__restrict__ active, int a0i,
                           const unsigned char *__restrict__ flags, const int *__restrict__ slot,
                           const double *__restrict__ mp_a, double *mp_b, double *mp_c) {
  extern __shared__ double sm[];
  double *sD = sm;                 // [12][nU]
  double *sa = sm + 12 * nU;       // [3*nmon]
  const int ntab = use_vs ? 12 : 3;
  if (G.dc) {
    if (g_stop(G)) return;
    a0i = G.dc[kGcActStart + lvl];
    if ((int)blockIdx.x >= G.dc[kGcActStart + lvl + 1] - a0i) return;
  }
  long g = active[a0i + blockIdx.x];
  long c = g - G.off[lvl];
  int ix, iy, iz;
  cell_decode(G, lvl, c, ix, iy, iz);
  int m = threadIdx.x;
  double b0 = 0, b1 = 0, b2 = 0;
  double cg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // enumerate the interaction list; all threads walk the same sequence
  int px_ = ix >> 1, py_ = iy >> 1, pz_ = iz >> 1;
  int total = (lvl == 0) ? G.ncl[0][0] * G.ncl[0][1] * G.ncl[0][2] : 27 * 8;
  for (int t = 0; t < total; ++t) {
    int sx, sy, sz;
    if (lvl == 0) {
      sx = t % G.ncl[0][0];
      sy = (t / G.ncl[0][0]) % G.ncl[0][1];
      sz = t / (G.ncl[0][0] * G.ncl[0][1]);
    } else {
      int nb = t >> 3, p = t & 7;
      int ic = nb % 3 - 1, jc = (nb / 3) % 3 - 1, kc = nb / 9 - 1;
      if (!ic && !jc && !kc) continue;
      int a = px_ + ic, b = py_ + jc, d = pz_ + kc;
      if (a < 0 || b < 0 || d < 0 || a >= G.ncl[lvl - 1][0] || b >= G.ncl[lvl - 1][1] ||
          d >= G.ncl[lvl - 1][2])
        continue;
      sx = 2 * a + (p & 1);
      sy = 2 * b + ((p >> 1) & 1);
      sz = 2 * d + ((p >> 2) & 1);
    }
    int di = sx - ix, dj = sy - iy, dk = sz - iz;
    if (abs(di) <= 1 && abs(dj) <= 1 && abs(dk) <= 1) continue;
    long gs = G.off[lvl] + cell_encode(G, lvl, sx, sy, sz);
    if (flags[gs] == 0) continue;
    const double *tab = &kd[kd_table(G, lvl, di, dj, dk) * 12 * (long)nU];
    const double *as = &mp_a[(long)slot[gs] * 3 * nmon];
    __syncthreads();
    for (int q = threadIdx.x; q < ntab * nU; q += blockDim.x) sD[q] = tab[q];
    for (int q = threadIdx.x; q < 3 * nmon; q += blockDim.x) sa[q] = as[q];
    __syncthreads();
    if (m < nmon) {
      double v0 = 0, v1 = 0, v2 = 0;
      double gg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int n = 0; n < nmon; ++n) {
        int id = mnT[n * nmon + m];
        double mu = multT[n * nmon + m];
        double A0 = sa[3 * n], A1 = sa[3 * n + 1], A2 = sa[3 * n + 2];
        double D0 = sD[id], D1 = sD[nU + id], D2 = sD[2 * nU + id];
        v0 += mu * (D1 * A2 - D2 * A1);
        v1 += mu * (D2 * A0 - D0 * A2);
        v2 += mu * (D0 * A1 - D1 * A0);
        if (use_vs) {
#pragma unroll
          for (int ii = 0; ii < 3; ++ii) {
            double E0 = sD[(3 + 0 + 3 * ii) * nU + id], E1 = sD[(3 + 1 + 3 * ii) * nU + id],
                   E2 = sD[(3 + 2 + 3 * ii) * nU + id];
            gg[ii + 0] += mu * (E1 * A2 - E2 * A1);
            gg[ii + 3] += mu * (E2 * A0 - E0 * A2);
            gg[ii + 6] += mu * (E0 * A1 - E1 * A0);
          }
        }
      }
      b0 += v0; b1 += v1; b2 += v2;
      for (int q = 0; q < 9; ++q) cg[q] += gg[q];
    }
  }
  if (m < nmon) {
    int i = pwr[3 * m], j = pwr[3 * m + 1], k = pwr[3 * m + 2];
    double sgn = ((i + j + k) & 1) ? -1.0 : 1.0;
    double *b = &mp_b[(long)slot[g] * 3 * nmon];
    b[3 * m] += sgn * b0; b[3 * m + 1] += sgn * b1; b[3 * m + 2] += sgn * b2;
    if (use_vs) {
      double *cc = &mp_c[(long)slot[g] * 9 * nmon];
      for (int q = 0; q < 9; ++q) cc[9 * m + q] += sgn * cg[q];
    }
  }
}

// L2L from level lvl to its active children (L2L_multipole,
// mod_multipole.f90:360-384):
// b_child(:,m) += sum_{s>=m} C(s,m) (c_child - c_parent)^(s-m) b_parent(:,s)
__global__ void k_l2l(TreeGeom G, int lvl, int deg, int nmon, int use_vs,
                      const int *__restrict__ pwr, const double *__restrict__ fact,
                      const int *__restrict__ active, int a0i, const int *__restrict__ slot,
                      double *mp_b, double *mp_c) {
  __shared__ double dp[3][11];
  __shared__ double sB[11 * 11];  // binomials fact[n] / (fact[r] fact[n-r]), as below
  if (G.dc) {
    if (g_stop(G)) return;
    a0i = G.dc[kGcActStart + lvl + 1];
    if ((int)blockIdx.x >= G.dc[kGcActStart + lvl + 2] - a0i) return;
  }
  for (int t = threadIdx.x; t < (deg + 1) * (deg + 1); t += blockDim.x) {
    const int n = t / (deg + 1), r = t - n * (deg + 1);
    sB[t] = r <= n ? fact[n] / (fact[r] * fact[n - r]) : 0.0;
  }
  long g = active[a0i + blockIdx.x];  // child at level lvl+1
  long c = g - G.off[lvl + 1];
  int ix, iy, iz;
  cell_decode(G, lvl + 1, c, ix, iy, iz);
  long gp = G.off[lvl] + (c >> 3);
  if (threadIdx.x < 3) {
    int d = threadIdx.x;
    int ci = d == 0 ? ix : (d == 1 ? iy : iz);
    double dd = cell_center(G, lvl + 1, d, ci) - cell_center(G, lvl, d, ci >> 1);
    double v = 1.0;
    for (int e = 0; e <= deg; ++e) { dp[d][e] = v; v *= dd; }
  }
  // the parent's local expansions (3 nmon + 9 nmon doubles), loaded together
  extern __shared__ double sP[];
  {
    const double *gb = &mp_b[(long)slot[gp] * 3 * nmon];
    const double *gc = &mp_c[(long)slot[gp] * 9 * nmon];
    for (int q = threadIdx.x; q < 3 * nmon; q += blockDim.x) sP[q] = gb[q];
    if (use_vs)
      for (int q = threadIdx.x; q < 9 * nmon; q += blockDim.x) sP[3 * nmon + q] = gc[q];
  }
  __syncthreads();
  int m = threadIdx.x;
  if (m >= nmon) return;
  int i = pwr[3 * m], j = pwr[3 * m + 1], k = pwr[3 * m + 2];
  const double *bp = sP;
  const double *cp = sP + 3 * nmon;
  double s0 = 0, s1 = 0, s2 = 0, sc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int ks = k; ks <= deg; ++ks)
    for (int js = j; js <= deg - ks; ++js)
      for (int is = i; is <= deg - ks - js; ++is) {
        int s = mono_index(is, js, ks);
        const int w = deg + 1;
        double nb = sB[is * w + i] * sB[js * w + j] * sB[ks * w + k];
        double f = nb * (dp[0][is - i] * dp[1][js - j] * dp[2][ks - k]);
        s0 += f * bp[3 * s];
        s1 += f * bp[3 * s + 1];
        s2 += f * bp[3 * s + 2];
        if (use_vs)
          for (int q = 0; q < 9; ++q) sc[q] += f * cp[9 * s + q];
      }
  double *b = &mp_b[(long)slot[g] * 3 * nmon];
  b[3 * m] += s0; b[3 * m + 1] += s1; b[3 * m + 2] += s2;
  if (use_vs) {
    double *cc = &mp_c[(long)slot[g] * 9 * nmon];
    for (int q = 0; q < 9; ++q) cc[9 * m + q] += sc[q];
  }
}

// ---------------------------------------------------------------------------
// Evaluation: L2P + P2P (apply_multipole, mod_octree.f90:1139-1429)
// ---------------------------------------------------------------------------

// Work units: every leaf's targets in chunks of <= 64 (see fmm_eval.h); each
// unit is cut into items of about seg_iters source iterations per lane.
__global__ void k_units(TreeGeom G, int n_leaves, const int *__restrict__ leaves,
                        const int *__restrict__ fstart, const int *__restrict__ tot, int seg_iters,
                        EvalUnit *units, int *n_units, int *n_items) {
  int lv = blockIdx.x * blockDim.x + threadIdx.x;
  if (G.dc) n_leaves = g_stop(G) ? 0 : G.dc[kGcLeaves];
  if (lv >= n_leaves) return;
  long g = leaves[lv];
  int lvl = lvl_of(G, g);
  int2 r = cell_range(G, fstart, lvl, g - G.off[lvl]);
  int np = r.y - r.x;
  if (np <= 0) return;
  const int T = tot[lv];
  int nu = (np + kWave - 1) / kWave;
  for (int q = 0; q < nu; ++q) {
    int nt = min(kWave, np - q * kWave);
    int lpt = 1;
    while (lpt * 2 * nt <= kWave) lpt *= 2;
    int iters = (T + lpt - 1) / lpt;
    int nseg = max(1, (iters + seg_iters - 1) / seg_iters);
    int i0 = atomicAdd(n_items, nseg);
    int w = atomicAdd(n_units, 1);
    if (G.dc && i0 + nseg > G.item_cap) atomicOr(&G.dc[kGcOvf], 8);
    units[w] = EvalUnit{lv, r.x + q * kWave, nt, lpt, i0, nseg, 0, 0};
  }
}

__global__ void k_items(int n_units, const EvalUnit *__restrict__ units,
                        const int *__restrict__ tot, EvalItem *items, const int *__restrict__ dc) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (dc) n_units = dc[kGcOvf] ? 0 : dc[kGcUnits];
  if (u >= n_units) return;
  EvalUnit U = units[u];
  int T = tot[U.leaf];
  int len = (T + U.nseg - 1) / U.nseg;
  for (int k = 0; k < U.nseg; ++k)
    items[U.item0 + k] = EvalItem{u, min(T, k * len), min(T, (k + 1) * len), 0};
}

// Leaf-averaged rate of strain of the L2P gradient (mod_octree.f90:1214-1224),
// needed for the turbulent viscosity; one wavefront per leaf.
__global__ void k_leaf_ros(EvalArgs A, int n_leaves) {
  extern __shared__ double sm[];
  const int nmon = A.nmon;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double *sb = sm + wv * 12 * nmon, *sc = sb + 3 * nmon;
  int *spw = (int *)(sm + kEvalWaves * 12 * nmon);
  for (int q = threadIdx.x; q < 3 * nmon; q += blockDim.x) spw[q] = A.pwr[q];
  __syncthreads();
  int lv = blockIdx.x * kEvalWaves + wv;
  if (A.G.dc) n_leaves = g_stop(A.G) ? 0 : A.G.dc[kGcLeaves];
  if (lv >= n_leaves) return;
  long g = A.leaves[lv];
  int lvl = lvl_of(A.G, g);
  long c = g - A.G.off[lvl];
  int ix, iy, iz;
  cell_decode(A.G, lvl, c, ix, iy, iz);
  double cx = cell_center(A.G, lvl, 0, ix), cy = cell_center(A.G, lvl, 1, iy),
         cz = cell_center(A.G, lvl, 2, iz);
  long sl = A.slot[g];
  for (int q = lane; q < 3 * nmon; q += kWave) sb[q] = A.mp_b[sl * 3 * nmon + q];
  for (int q = lane; q < 9 * nmon; q += kWave) sc[q] = A.mp_c[sl * 9 * nmon + q];
  wave_sync();
  int2 own = cell_range(A.G, A.fstart, lvl, c);
  double acc = 0.0;
  for (int t = own.x + lane; t < own.y; t += kWave) {
    double v[3], gr[9];
    l2p_point(A.deg, nmon, 1, sb, sc, spw, A.px[t] - cx, A.py[t] - cy, A.pz[t] - cz, v, gr);
    for (int ii = 0; ii < 3; ++ii)
      for (int jj = 0; jj < 3; ++jj) {
        double e = 0.5 * (gr[ii + 3 * jj] + gr[jj + 3 * ii]);
        acc += e * e;
      }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
  if (lane == 0) {
    int np = own.y - own.x;
    const_cast<double *>(A.leaf_ros)[lv] = np > 0 ? acc / (double)np : 0.0;
  }
}

__device__ inline double turbvisc_of(const EvalArgs &A, int leaf, int t) {
  if (!A.use_tv) return 0.0;
  double rr = 0.11 * A.pr[t];
  return rr * rr * sqrt(2.0 * A.leaf_ros[leaf]);
}

// P2P of one item, plain FP64.  Pair terms are the reference formulas
// (compute_vel, compute_stretch, compute_diffusion of mod_vortpart.f90) in
// FP64; the list is walked in the reference order (neighbours, leaf_neigh,
// own leaf).  Writes the item's partial sums (velocity and stretching,
// divided by 4 pi; diffusion times (nu + nu_t)).
__global__ void __launch_bounds__(kWave * kEvalWaves) k_p2p_fp64(EvalArgs A) {
  __shared__ double ssm[kEvalWaves][8 * kWave];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double *ss = ssm[wv];
  const int it = blockIdx.x * kEvalWaves + wv;
  if (it >= (A.G.dc ? (g_stop(A.G) ? 0 : A.G.dc[kGcItems]) : A.n_items)) return;
  const EvalItem I = A.items[it];
  const EvalUnit U = A.units[I.unit];
  const int lpt = U.lpt, tl = lane / lpt, sub = lane - tl * lpt;
  const bool valid = tl < U.nt;
  const int t = U.t0 + (valid ? tl : 0);
  const double x = A.px[t], y = A.py[t], z = A.pz[t];
  const double mg = A.pm[t];
  const double al0 = mg * A.dx[t], al1 = mg * A.dy[t], al2 = mg * A.dz[t];
  const double tvp = turbvisc_of(A, U.leaf, t);
  double pv0 = 0, pv1 = 0, pv2 = 0, ps0 = 0, ps1 = 0, ps2 = 0, pd0 = 0, pd1 = 0, pd2 = 0;
  double pr0 = 0, pr1 = 0, pr2 = 0;
  const int nr = A.cnt[U.leaf];
  const int2 *rl = &A.rng[(long)U.leaf * A.cap];
  int ri, off;
  seg_seek(rl, nr, I.lo, ri, off);
  int left = I.hi - I.lo;
  for (; ri < nr && left > 0; ++ri, off = 0) {
    const int2 r = rl[ri];
    const int e = min(r.y, r.x + off + left);
    for (int s0 = r.x + off; s0 < e; s0 += kWave) {
      const int ns = min(kWave, e - s0);
      if (lane < ns) {
        const int q = s0 + lane;
        ss[0 * kWave + lane] = A.px[q];
        ss[1 * kWave + lane] = A.py[q];
        ss[2 * kWave + lane] = A.pz[q];
        ss[3 * kWave + lane] = A.dx[q];
        ss[4 * kWave + lane] = A.dy[q];
        ss[5 * kWave + lane] = A.dz[q];
        ss[6 * kWave + lane] = A.pm[q];
        ss[7 * kWave + lane] = A.pr[q];
      }
      wave_sync();
      if (valid) {
        for (int s = sub; s < ns; s += lpt) {
          if (s0 + s == t) continue;
          double r0 = x - ss[s], r1 = y - ss[kWave + s], r2 = z - ss[2 * kWave + s];
          double sdx = ss[3 * kWave + s], sdy = ss[4 * kWave + s], sdz = ss[5 * kWave + s];
          double smg = ss[6 * kWave + s], srv = ss[7 * kWave + s];
          double d2 = r0 * r0 + r1 * r1 + r2 * r2;
          // kernel_coeffs (mod_vortpart.f90:269-271; Gaussian: 283-291)
          double cf, df;
          if (A.kernel == DUST_KERNEL_GAUSSIAN) {
            gauss_cd(d2, srv, cf, df);
          } else {
            double R2 = d2 + srv * srv;
            double R3 = R2 * sqrt(R2);
            cf = -1.0 / R3;
            df = 3.0 / (R3 * R2);
          }
          pv0 += cf * (r1 * sdz - r2 * sdy) * smg;
          pv1 += cf * (r2 * sdx - r0 * sdz) * smg;
          pv2 += cf * (r0 * sdy - r1 * sdx) * smg;
          if (A.use_vs) {
            double am0 = sdx * smg, am1 = sdy * smg, am2 = sdz * smg;
            double vp0 = al1 * am2 - al2 * am1, vp1 = al2 * am0 - al0 * am2, vp2 = al0 * am1 - al1 * am0;
            double dv = r0 * vp0 + r1 * vp1 + r2 * vp2;
            ps0 -= cf * vp0 + df * r0 * dv;
            ps1 -= cf * vp1 + df * r1 * dv;
            ps2 -= cf * vp2 + df * r2 * dv;
            if (A.divf) {
              // compute_rotu (mod_vortpart.f90:236-256): -2 c alpha_s + d r x (r x alpha_s)
              const double i0 = r1 * am2 - r2 * am1, i1 = r2 * am0 - r0 * am2,
                           i2 = r0 * am1 - r1 * am0;
              pr0 += -2.0 * cf * sdx * smg + df * (r1 * i2 - r2 * i1);
              pr1 += -2.0 * cf * sdy * smg + df * (r2 * i0 - r0 * i2);
              pr2 += -2.0 * cf * sdz * smg + df * (r0 * i1 - r1 * i0);
            }
          }
          if (A.use_vd) {
            // compute_diffusion with etaeps (mod_vortpart.f90:305-336)
            double vol = 4.0 / 3.0 * kPi * srv * srv * srv;
            double qq = d2 / (srv * srv) + 1.0;
            double q2 = qq * qq;
            double eta = 105.0 / (8.0 * kPi) / (q2 * q2 * sqrt(qq)) / (srv * srv * srv);
            double f = 1.0 / (srv * srv) * eta * (A.nu_inf + tvp);
            pd0 += f * (vol * sdx * smg - vol * al0);
            pd1 += f * (vol * sdy * smg - vol * al1);
            pd2 += f * (vol * sdz * smg - vol * al2);
          }
        }
      }
      wave_sync();
    }
    left -= e - (r.x + off);
  }
  pv0 = lane_group_sum(pv0, lpt); pv1 = lane_group_sum(pv1, lpt); pv2 = lane_group_sum(pv2, lpt);
  if (A.use_vs) {
    ps0 = lane_group_sum(ps0, lpt); ps1 = lane_group_sum(ps1, lpt); ps2 = lane_group_sum(ps2, lpt);
  }
  if (A.divf) {
    pr0 = lane_group_sum(pr0, lpt); pr1 = lane_group_sum(pr1, lpt); pr2 = lane_group_sum(pr2, lpt);
  }
  if (A.use_vd) {
    pd0 = lane_group_sum(pd0, lpt); pd1 = lane_group_sum(pd1, lpt); pd2 = lane_group_sum(pd2, lpt);
  }
  if (valid && sub == 0) {
    const double fi = 1.0 / (4.0 * kPi);
    double *o = &A.part[((long)it * kWave + tl) * A.pstride];
    o[0] = pv0 * fi; o[1] = pv1 * fi; o[2] = pv2 * fi;
    o[3] = ps0 * fi + pd0; o[4] = ps1 * fi + pd1; o[5] = ps2 * fi + pd2;
    if (A.divf) {
      o[6] = pr0 * fi; o[7] = pr1 * fi; o[8] = pr2 * fi;
    }
  }
}

// Finish of every unit (both forms): L2P in FP64, the items' P2P sums in
// item order, the free stream, and the turbulent viscosity.
__global__ void __launch_bounds__(kWave * kEvalWaves) k_eval_finish(EvalArgs A) {
  extern __shared__ double sm[];
  const int nmon = A.nmon;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double *sb = sm + wv * 12 * nmon, *sc = sb + 3 * nmon;
  int *spw = (int *)(sm + kEvalWaves * 12 * nmon);
  for (int q = threadIdx.x; q < 3 * nmon; q += blockDim.x) spw[q] = A.pwr[q];
  __syncthreads();
  const int u = blockIdx.x * kEvalWaves + wv;
  if (u >= (A.G.dc ? (g_stop(A.G) ? 0 : A.G.dc[kGcUnits]) : A.n_units)) return;
  const EvalUnit U = A.units[u];
  const long g = A.leaves[U.leaf];
  const int lvl = lvl_of(A.G, g);
  const long c = g - A.G.off[lvl];
  int ix, iy, iz;
  cell_decode(A.G, lvl, c, ix, iy, iz);
  const double cx = cell_center(A.G, lvl, 0, ix), cy = cell_center(A.G, lvl, 1, iy),
               cz = cell_center(A.G, lvl, 2, iz);
  const long sl = A.slot[g];
  for (int q = lane; q < 3 * nmon; q += kWave) sb[q] = A.mp_b[sl * 3 * nmon + q];
  if (A.use_vs)
    for (int q = lane; q < 9 * nmon; q += kWave) sc[q] = A.mp_c[sl * 9 * nmon + q];
  wave_sync();
  if (lane >= U.nt) return;
  const int t = U.t0 + lane;
  const double x = A.px[t], y = A.py[t], z = A.pz[t];
  const double mg = A.pm[t];
  const double al0 = mg * A.dx[t], al1 = mg * A.dy[t], al2 = mg * A.dz[t];
  double v[3], gr[9], st[3] = {0, 0, 0};
  l2p_point(A.deg, nmon, A.use_vs, sb, sc, spw, x - cx, y - cy, z - cz, v, gr);
  if (A.use_vs)  // str = matmul(alpha, transpose(grad)): str_j = sum_i grad(j,i) alpha_i
    for (int jj = 0; jj < 3; ++jj) st[jj] = gr[jj] * al0 + gr[jj + 3] * al1 + gr[jj + 6] * al2;
  // partial sums of the unit's items, added in item order; the loads of four
  // items are issued together (units of large leaves have hundreds of items)
  double pv[3] = {0, 0, 0}, ps[3] = {0, 0, 0};
  const int ps_ = A.pstride;
  const double *o = &A.part[((long)U.item0 * kWave + lane) * ps_];
  int k = 0;
  for (; k + 4 <= U.nseg; k += 4) {
    double q[4][6];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int e = 0; e < 6; ++e) q[u][e] = o[(long)(k + u) * kWave * ps_ + e];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        pv[e] += q[u][e];
        ps[e] += q[u][3 + e];
      }
  }
  for (; k < U.nseg; ++k)
    for (int e = 0; e < 3; ++e) {
      pv[e] += o[(long)k * kWave * ps_ + e];
      ps[e] += o[(long)k * kWave * ps_ + 3 + e];
    }
  // the wind at the particle (mod_octree.f90:1376): u_inf without a gust
  double wnd[3];
  const double pos[3] = {x, y, z};
  variable_wind(A.wind, pos, A.time, wnd);
  for (int e = 0; e < 3; ++e) {
    A.vel[3 * t + e] = v[e] + pv[e] + wnd[e];
    A.str[3 * t + e] = st[e] + ps[e];
  }
  if (A.divf) {
    // rotu = curl of the local expansion's velocity, grad(2,3) - grad(3,2),
    // ... (mod_octree.f90:1207-1212; gr[i + 3 j] = grad(i+1, j+1)), then
    // the items' P2P sums in item order
    double ru[3] = {gr[7] - gr[5], gr[2] - gr[6], gr[3] - gr[1]};
    for (int q = 0; q < U.nseg; ++q)
      for (int e = 0; e < 3; ++e) ru[e] += o[(long)q * kWave * ps_ + 6 + e];
    for (int e = 0; e < 3; ++e) A.rot[3 * t + e] = ru[e];
  }
  A.tv[t] = turbvisc_of(A, U.leaf, t);
}

size_t eval_finish_lds(int nmon) {
  return sizeof(double) * kEvalWaves * 12 * nmon + sizeof(int) * 3 * nmon;
}

// Scatter results from sorted to input order.
__global__ void k_scatter(int n, const int *__restrict__ perm, const double *__restrict__ vs,
                          const double *__restrict__ ss, const double *__restrict__ ts,
                          double *v, double *s, double *tv, const int *__restrict__ n_dev) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (n_dev) n = *n_dev;
  if (i >= n) return;
  int q = perm[i];
  v[3 * q] = vs[3 * i];
  v[3 * q + 1] = vs[3 * i + 1];
  v[3 * q + 2] = vs[3 * i + 2];
  s[3 * q] = ss[3 * i];
  s[3 * q + 1] = ss[3 * i + 1];
  s[3 * q + 2] = ss[3 * i + 2];
  tv[q] = ts[i];
}

// Scatter a 3-vector per particle from sorted to input order.
__global__ void k_scatter3(int n, const int *__restrict__ perm, const double *__restrict__ src,
                           double *dst, const int *__restrict__ n_dev) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (n_dev) n = *n_dev;
  if (i >= n) return;
  int q = perm[i];
  dst[3 * q] = src[3 * i];
  dst[3 * q + 1] = src[3 * i + 1];
  dst[3 * q + 2] = src[3 * i + 2];
}

constexpr int kEvalThreads = 128;

// Direct O(N^2) particle terms, the fmm=F branch (mod_wake.f90:1050-1113).
__global__ void __launch_bounds__(kEvalThreads)
    k_direct(int n, const double *__restrict__ cen, const double *__restrict__ dir,
             const double *__restrict__ mag, const double *__restrict__ rv, int use_vs, int use_vd,
             double nu_inf, double u0, double u1, double u2, int kernel, double *vel, double *str) {
  __shared__ double ss[8][kEvalThreads];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = t < n;
  double x = 0, y = 0, z = 0, al0 = 0, al1 = 0, al2 = 0;
  if (valid) {
    x = cen[3 * t]; y = cen[3 * t + 1]; z = cen[3 * t + 2];
    al0 = dir[3 * t] * mag[t]; al1 = dir[3 * t + 1] * mag[t]; al2 = dir[3 * t + 2] * mag[t];
  }
  double pv0 = 0, pv1 = 0, pv2 = 0, ps0 = 0, ps1 = 0, ps2 = 0, pd0 = 0, pd1 = 0, pd2 = 0;
  for (int b0 = 0; b0 < n; b0 += kEvalThreads) {
    int ns = min(kEvalThreads, n - b0);
    __syncthreads();
    if (threadIdx.x < ns) {
      int q = b0 + threadIdx.x;
      ss[0][threadIdx.x] = cen[3 * q];
      ss[1][threadIdx.x] = cen[3 * q + 1];
      ss[2][threadIdx.x] = cen[3 * q + 2];
      ss[3][threadIdx.x] = dir[3 * q];
      ss[4][threadIdx.x] = dir[3 * q + 1];
      ss[5][threadIdx.x] = dir[3 * q + 2];
      ss[6][threadIdx.x] = mag[q];
      ss[7][threadIdx.x] = rv[q];
    }
    __syncthreads();
    if (!valid) continue;
    for (int s = 0; s < ns; ++s) {
      if (b0 + s == t) continue;
      double r0 = x - ss[0][s], r1 = y - ss[1][s], r2 = z - ss[2][s];
      double sdx = ss[3][s], sdy = ss[4][s], sdz = ss[5][s], smg = ss[6][s], srv = ss[7][s];
      double d2 = r0 * r0 + r1 * r1 + r2 * r2;
      double cf, df;
      if (kernel == DUST_KERNEL_GAUSSIAN) {
        gauss_cd(d2, srv, cf, df);
      } else {
        double R2 = d2 + srv * srv;
        double R3 = R2 * sqrt(R2);
        cf = -1.0 / R3;
        df = 3.0 / (R3 * R2);
      }
      pv0 += cf * (r1 * sdz - r2 * sdy) * smg;
      pv1 += cf * (r2 * sdx - r0 * sdz) * smg;
      pv2 += cf * (r0 * sdy - r1 * sdx) * smg;
      if (use_vs) {
        double am0 = sdx * smg, am1 = sdy * smg, am2 = sdz * smg;
        double vp0 = al1 * am2 - al2 * am1, vp1 = al2 * am0 - al0 * am2, vp2 = al0 * am1 - al1 * am0;
        double dv = r0 * vp0 + r1 * vp1 + r2 * vp2;
        ps0 -= cf * vp0 + df * r0 * dv;
        ps1 -= cf * vp1 + df * r1 * dv;
        ps2 -= cf * vp2 + df * r2 * dv;
      }
      if (use_vd) {
        double vol = 4.0 / 3.0 * kPi * srv * srv * srv;
        double qq = d2 / (srv * srv) + 1.0;
        double q2 = qq * qq;
        double eta = 105.0 / (8.0 * kPi) / (q2 * q2 * sqrt(qq)) / (srv * srv * srv);
        double f = 1.0 / (srv * srv) * eta * nu_inf;
        pd0 += f * (vol * sdx * smg - vol * al0);
        pd1 += f * (vol * sdy * smg - vol * al1);
        pd2 += f * (vol * sdz * smg - vol * al2);
      }
    }
  }
  if (valid) {
    const double fi = 1.0 / (4.0 * kPi);
    vel[3 * t] = pv0 * fi + u0;
    vel[3 * t + 1] = pv1 * fi + u1;
    vel[3 * t + 2] = pv2 * fi + u2;
    if (str) {
      str[3 * t] = ps0 * fi + pd0;
      str[3 * t + 1] = ps1 * fi + pd1;
      str[3 * t + 2] = ps2 * fi + pd2;
    }
  }
}

// ---------------------------------------------------------------------------
// Stage timing
// ---------------------------------------------------------------------------
// Events come from a pool that is kept between passes (no event creation in
// a timed loop); marks index pairs of pool events.
void timer_begin(FmmHandle *h, int stage) {
  if (!h->timer.on) return;
  const size_t need = 2 * (h->timer.marks.size() + 1);
  while (h->timer.pool.size() < need) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    h->timer.pool.push_back(e);
  }
  const int i = (int)need - 2;
  h->timer.marks.push_back({stage, i});
  (void)hipEventRecord(h->timer.pool[i], h->stream);
}
void timer_end(FmmHandle *h, int stage) {
  if (!h->timer.on || h->timer.marks.empty()) return;
  (void)stage;
  (void)hipEventRecord(h->timer.pool[h->timer.marks.back().second + 1], h->stream);
}
static void timer_collect(FmmHandle *h) {
  if (h->timer.marks.empty()) return;
  (void)hipStreamSynchronize(h->stream);
  for (auto &mk : h->timer.marks) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->timer.pool[mk.second], h->timer.pool[mk.second + 1]) ==
        hipSuccess) {
      h->timer.ms[mk.first] += ms;
      h->timer.launches[mk.first] += 1;
    }
  }
  h->timer.marks.clear();
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
static int fmm_create(const dust_fmm_params *prm, int capacity, FmmHandle **out) {
  DUST_CHECK(prm && out, DUST_E_ARG, "null argument");
  DUST_CHECK(prm->nlevels >= 1 && prm->nlevels <= kMaxLevels, DUST_E_ARG,
             "nlevels must be in 1..%d", kMaxLevels);
  DUST_CHECK(prm->degree >= 0 && prm->degree <= 8, DUST_E_ARG, "degree must be in 0..8");
  DUST_CHECK(prm->nbox[0] >= 1 && prm->nbox[1] >= 1 && prm->nbox[2] >= 1, DUST_E_ARG,
             "nbox must be positive");
  DUST_CHECK(prm->box_length > 0.0, DUST_E_ARG, "box_length must be positive");
  DUST_CHECK(capacity >= 1, DUST_E_ARG, "capacity must be positive");
  DUST_CHECK(prm->precision == DUST_FMM_FP64 || prm->precision == DUST_FMM_TUNED, DUST_E_ARG,
             "unknown precision %d", prm->precision);
  DUST_CHECK(prm->kernel == DUST_KERNEL_ROSENHEAD || prm->kernel == DUST_KERNEL_GAUSSIAN, DUST_E_ARG,
             "kernel must be DUST_KERNEL_ROSENHEAD or DUST_KERNEL_GAUSSIAN");
  long nb = (long)prm->nbox[0] * prm->nbox[1] * prm->nbox[2];
  long nfine = nb << (3 * (prm->nlevels - 1));
  DUST_CHECK(nfine < (1L << 31) - 1, DUST_E_ARG, "finest level has too many cells");
  FmmHandle *h = new FmmHandle();
  h->prm = *prm;
  for (int e = 0; e < 3; ++e) h->wind.u_inf[e] = prm->u_inf[e];
  h->wind.type = DUST_GUST_NONE;
  h->cap = capacity;
  h->nlev = prm->nlevels;
  h->deg = prm->degree;
  h->nmon = mono_count(h->deg);
  h->nU = mono_count(2 * h->deg);
  for (int d = 0; d < 3; ++d) {
    h->xmin[d] = prm->origin[d];
    h->xmax[d] = prm->origin[d] + (double)prm->nbox[d] * prm->box_length;
  }
  long off = 0, kdb = 0;
  for (int l = 0; l < h->nlev; ++l) {
    Level &L = h->lev[l];
    for (int d = 0; d < 3; ++d) {
      L.ncl[d] = prm->nbox[d] << l;
      L.kdl[d] = (l == 0) ? std::max(prm->nbox[d], 3) : 3;
    }
    L.ncells = (long)L.ncl[0] * L.ncl[1] * L.ncl[2];
    L.off = off;
    off += L.ncells;
    L.csize = prm->box_length / (double)(1L << l);
    L.kd_base = kdb;
    kdb += (long)(2 * L.kdl[0] + 1) * (2 * L.kdl[1] + 1) * (2 * L.kdl[2] + 1);
  }
  h->ncells_tot = off;
  h->nfine = h->lev[h->nlev - 1].ncells;
  h->kd_tables = kdb;
  int rc = 0;
  auto fail = [&](int code) {
    delete h;
    return code;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(DUST_E_HIP);
  }
  if (h->rb.reserve(1)) return fail(DUST_E_NOMEM);
  // polynomial tables
  std::vector<int> pw;
  mono_powers(h->deg, pw);
  int nm = h->nmon;
  std::vector<double> multT((size_t)nm * nm);
  std::vector<int> mnT((size_t)nm * nm);
  std::vector<double> fact(2 * h->deg + 3);
  fact[0] = 1.0;
  for (size_t q = 1; q < fact.size(); ++q) fact[q] = fact[q - 1] * (double)q;
  for (int m = 0; m < nm; ++m)
    for (int n = 0; n < nm; ++n) {
      int a = pw[3 * m] + pw[3 * n], b = pw[3 * m + 1] + pw[3 * n + 1], c = pw[3 * m + 2] + pw[3 * n + 2];
      // mult = nfact_der(m+n)/(nfact(m) nfact(n)), mod_multipole.f90:304-306,
      // with exact factorials
      multT[(size_t)n * nm + m] = (fact[a] * fact[b] * fact[c]) /
                                  ((fact[pw[3 * m]] * fact[pw[3 * m + 1]] * fact[pw[3 * m + 2]]) *
                                   (fact[pw[3 * n]] * fact[pw[3 * n + 1]] * fact[pw[3 * n + 2]]));
      mnT[(size_t)n * nm + m] = mono_index(a, b, c);
    }
  // kernel derivative tables, initialize_octree (mod_octree.f90:486-516):
  // offset (i,j,k) has displacement cell_size * (-i,-j,-k)
  h->h_kd.assign((size_t)kdb * 12 * h->nU, 0.0);
  for (int l = 0; l < h->nlev; ++l) {
    Level &L = h->lev[l];
    int w0 = 2 * L.kdl[0] + 1, w1 = 2 * L.kdl[1] + 1;
    for (int k = -L.kdl[2]; k <= L.kdl[2]; ++k)
      for (int j = -L.kdl[1]; j <= L.kdl[1]; ++j)
        for (int i = -L.kdl[0]; i <= L.kdl[0]; ++i) {
          if (std::abs(i) < 2 && std::abs(j) < 2 && std::abs(k) < 2) continue;
          long o = L.kd_base + (i + L.kdl[0]) + w0 * ((j + L.kdl[1]) + (long)w1 * (k + L.kdl[2]));
          double diff[3] = {L.csize * (double)(-i), L.csize * (double)(-j), L.csize * (double)(-k)};
          kernel_derivatives(h->deg, diff, prm->delta, h->nU, &h->h_kd[(size_t)o * 12 * h->nU]);
        }
  }
  if ((rc = h->d_pwr.reserve(pw.size())) || (rc = h->d_multT.reserve(multT.size())) ||
      (rc = h->d_mnT.reserve(mnT.size())) || (rc = h->d_kd.reserve(h->h_kd.size())) ||
      (rc = h->d_fact.reserve(fact.size())))
    return fail(rc);
  if (hipMemcpy(h->d_pwr.p, pw.data(), pw.size() * sizeof(int), hipMemcpyHostToDevice) ||
      hipMemcpy(h->d_multT.p, multT.data(), multT.size() * sizeof(double), hipMemcpyHostToDevice) ||
      hipMemcpy(h->d_mnT.p, mnT.data(), mnT.size() * sizeof(int), hipMemcpyHostToDevice) ||
      hipMemcpy(h->d_kd.p, h->h_kd.data(), h->h_kd.size() * sizeof(double), hipMemcpyHostToDevice) ||
      hipMemcpy(h->d_fact.p, fact.data(), fact.size() * sizeof(double), hipMemcpyHostToDevice)) {
    set_error("table upload failed");
    return fail(DUST_E_HIP);
  }
  // particle and tree storage sized for the capacity
  size_t C = (size_t)capacity;
  if ((rc = h->cen_in.reserve(3 * C)) || (rc = h->dir_in.reserve(3 * C)) ||
      (rc = h->mag_in.reserve(C)) || (rc = h->rv_in.reserve(C)) || (rc = h->key.reserve(C)) ||
      (rc = h->key_s.reserve(C)) || (rc = h->idx_in.reserve(C)) || (rc = h->perm.reserve(C)) ||
      (rc = h->px.reserve(C)) || (rc = h->py.reserve(C)) || (rc = h->pz.reserve(C)) ||
      (rc = h->dx.reserve(C)) || (rc = h->dy.reserve(C)) || (rc = h->dz.reserve(C)) ||
      (rc = h->pm.reserve(C)) || (rc = h->pr.reserve(C)) || (rc = h->vel_s.reserve(3 * C)) ||
      (rc = h->str_s.reserve(3 * C)) || (rc = h->tv_s.reserve(C)) || (rc = h->vel.reserve(3 * C)) ||
      (rc = h->str.reserve(3 * C)) || (rc = h->tv.reserve(C)) ||
      (rc = h->fstart.reserve(h->nfine + 1)) || (rc = h->flags.reserve(h->ncells_tot)) ||
      (rc = h->slot.reserve(h->ncells_tot)) || (rc = h->active.reserve(h->ncells_tot)) ||
      (rc = h->leaves.reserve(h->ncells_tot)) || (rc = h->counts.reserve(kMaxLevels + 4)) ||
      (rc = h->d_err.reserve(1)) || (rc = h->n_units.reserve(2)))
    return fail(rc);
  if (prm->use_vs && prm->use_divfilt &&
      ((rc = h->rot_s.reserve(3 * C)) || (rc = h->rot.reserve(3 * C))))
    return fail(rc);
  // scratch for the device-wide primitives, sized once so that later calls
  // never reallocate while work is in flight
  if ((rc = prim_reserve(h))) return fail(rc);
  {
    size_t m2l_lds = sizeof(double) * (12 * h->nU + 3 * nm);
    if (m2l_lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)k_m2l_fp64, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)m2l_lds) != hipSuccess) {
      set_error("cannot reserve %zu bytes of LDS for M2L", m2l_lds);
      return fail(DUST_E_HIP);
    }
  }
  {
    size_t ev = eval_finish_lds(nm);
    if (ev > 64 * 1024 &&
        hipFuncSetAttribute((const void *)k_eval_finish, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)ev) != hipSuccess) {
      set_error("cannot reserve %zu bytes of LDS for the evaluation", ev);
      return fail(DUST_E_HIP);
    }
  }
  h->p2p_cap = 27 + 26 * (h->nlev - 1);
  if (prm->precision == DUST_FMM_TUNED && (rc = tuned_prepare(h))) return fail(rc);
  *out = h;
  return 0;
}

static void fmm_destroy(FmmHandle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  tuned_release(h);
  DevBuf<double> *dbl[] = {&h->d_multT, &h->d_kd, &h->d_fact, &h->cen_in, &h->dir_in, &h->mag_in,
                           &h->rv_in, &h->px, &h->py, &h->pz, &h->dx, &h->dy, &h->dz, &h->pm,
                           &h->pr, &h->mp_a, &h->mp_b, &h->mp_c, &h->vel_s, &h->str_s, &h->tv_s,
                           &h->vel, &h->str, &h->tv, &h->rot_s, &h->rot};
  for (auto b : dbl) b->release();
  DevBuf<int> *ints[] = {&h->d_pwr, &h->d_mnT, &h->idx_in, &h->perm, &h->fstart, &h->slot,
                         &h->active, &h->leaves, &h->counts, &h->p2p_cnt, &h->d_err,
                         &h->n_units};
  h->units.release();
  h->items.release();
  h->part.release();
  h->p2p_tot.release();
  h->leaf_ros.release();
  for (auto b : ints) b->release();
  h->key.release();
  h->key_s.release();
  h->tmp.release();
  h->flags.release();
  h->p2p_rng.release();
  h->rb.release();
  DevBuf<double> *prd[] = {&h->pr_cen, &h->pr_dir, &h->pr_mag};
  for (auto b : prd) b->release();
  for (auto &b : h->pr_stage) b.release();
  DevBuf<unsigned char> *pru[] = {&h->pr_has, &h->pr_near, &h->pr_keep_s, &h->pr_keep};
  for (auto b : pru) b->release();
  DevBuf<int> *pri[] = {&h->pr_sel, &h->pr_ns, &h->pr_perm};
  for (auto b : pri) b->release();
  h->pr_key.release();
  if (h->gr.exec) (void)hipGraphExecDestroy(h->gr.exec);
  if (h->gr.graph) (void)hipGraphDestroy(h->gr.graph);
  h->gr.dc.release();
  h->gr.rs_hist.release();
  h->gr.sel_cnt.release();
  for (auto e : h->timer.pool) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

static int fmm_set_particles(FmmHandle *h, int n, const double *cen, const double *dir,
                             const double *mag, const double *rv, hipMemcpyKind kind) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  DUST_CHECK(n >= 0, DUST_E_ARG, "negative particle count");
  DUST_CHECK(n <= h->cap, DUST_E_CAPACITY, "%d particles exceed the handle capacity %d", n, h->cap);
  DUST_CHECK(n == 0 || (cen && dir && mag && rv), DUST_E_ARG, "null particle array");
  h->n = n;
  h->sorted = h->computed = h->applied = false;
  if (n == 0) return 0;
  DUST_HIP(hipMemcpyAsync(h->cen_in.p, cen, sizeof(double) * 3 * n, kind, h->stream));
  DUST_HIP(hipMemcpyAsync(h->dir_in.p, dir, sizeof(double) * 3 * n, kind, h->stream));
  DUST_HIP(hipMemcpyAsync(h->mag_in.p, mag, sizeof(double) * n, kind, h->stream));
  DUST_HIP(hipMemcpyAsync(h->rv_in.p, rv, sizeof(double) * n, kind, h->stream));
  if (kind == hipMemcpyHostToDevice) DUST_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

// The redistribution block of sort_particles (mod_octree.f90:827-894) on the
// sorted particles, with fstart over all of them: near-wall marks, the
// per-cell loop, the outcome in input order, then the kept particles
// compacted to the front of the sorted arrays (their order, hence the push
// order of every cell, unchanged) and fstart recomputed over them.  The kept
// count stays on the device (pr_ns) until the phase's read-back.
static int redistribute(FmmHandle *h) {
  const int n = h->n;
  TreeGeom G = make_geom(h);
  const int TB = 256;
  const int ls = std::min(std::max(h->pr_lvl, 0), h->nlev - 1);
  const long nls = h->lev[ls].ncells;
  DUST_TRY(h->pr_has.reserve(nls));
  DUST_TRY(h->pr_near.reserve(nls));
  DUST_TRY(h->pr_keep_s.reserve(h->cap));
  DUST_TRY(h->pr_keep.reserve(h->cap));
  DUST_TRY(h->pr_dir.reserve((size_t)3 * h->cap));
  DUST_TRY(h->pr_mag.reserve(h->cap));
  DUST_TRY(h->pr_sel.reserve(h->cap));
  DUST_TRY(h->pr_ns.reserve(1));
  DUST_TRY(h->pr_perm.reserve(h->cap));
  DUST_TRY(h->pr_key.reserve(h->cap));
  for (auto &b : h->pr_stage) DUST_TRY(b.reserve(h->cap));
  k_fill_u8<<<ceil_div(nls, TB), TB, 0, h->stream>>>(nls, h->pr_has.p, 0);
  if (h->pr_m > 0)
    k_pr_bin<<<ceil_div(h->pr_m, TB), TB, 0, h->stream>>>(G, h->pr_m, h->pr_cen.p, ls,
                                                          h->pr_has.p);
  k_pr_near<<<ceil_div(nls, TB), TB, 0, h->stream>>>(G, ls, nls, h->pr_has.p, h->pr_near.p);
  k_pr_cells<<<ceil_div(h->nfine, TB), TB, 0, h->stream>>>(G, ls, h->fstart.p, h->pr_near.p,
                                                           1.0 / h->pr_ratio, h->pm.p, h->dx.p,
                                                           h->dy.p, h->dz.p, h->pr_keep_s.p);
  k_pr_out<<<ceil_div(n, TB), TB, 0, h->stream>>>(n, h->perm.p, h->pr_keep_s.p, h->pm.p, h->dx.p,
                                                  h->dy.p, h->dz.p, h->pr_keep.p, h->pr_dir.p,
                                                  h->pr_mag.p);
  DUST_TRY(prim_select_kept(h, h->pr_keep_s.p, n, h->pr_sel.p, h->pr_ns.p));
  DevBuf<double> *arr[8] = {&h->px, &h->py, &h->pz, &h->dx, &h->dy, &h->dz, &h->pm, &h->pr};
  PrArrays A;
  for (int a = 0; a < 8; ++a) {
    A.src[a] = arr[a]->p;
    A.dst[a] = h->pr_stage[a].p;
  }
  k_pr_compact<<<ceil_div(n, TB), TB, 0, h->stream>>>(n, h->pr_ns.p, h->pr_sel.p, A, h->key_s.p,
                                                      h->pr_key.p, h->perm.p, h->pr_perm.p);
  for (int a = 0; a < 8; ++a) std::swap(*arr[a], h->pr_stage[a]);
  std::swap(h->key_s, h->pr_key);
  std::swap(h->perm, h->pr_perm);
  k_fstart<<<ceil_div(h->nfine + 1, TB), TB, 0, h->stream>>>(h->nfine, n, h->pr_ns.p, h->key_s.p,
                                                             h->fstart.p);
  h->pr_done = true;
  return 0;
}

// sort_particles (mod_octree.f90:703-996)
// sort_particles in three phases separated by host read-backs of device
// counts (fmm_phase_* below); the read-back targets live in the handle so
// that several handles can be advanced between synchronisations.
static int sort_phase_a(FmmHandle *h) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  const int n = h->n;
  TreeGeom G = make_geom(h);
  const int TB = 256;
  timer_begin(h, DUST_FMM_STAGE_SORT);
  FmmReadback *rb = h->rb.p;
  rb->init_err = 0x7fffffff;
  DUST_HIP(hipMemcpyAsync(h->d_err.p, &rb->init_err, sizeof(int), hipMemcpyHostToDevice,
                          h->stream));
  if (n > 0) {
    k_keys<<<ceil_div(n, TB), TB, 0, h->stream>>>(G, n, h->cen_in.p, h->key.p, h->idx_in.p,
                                                  h->d_err.p);
    int bits = 0;
    while ((1L << bits) < h->nfine) ++bits;
    if (bits == 0) bits = 1;
    DUST_TRY(prim_sort_pairs(h, n, bits));
    k_gather<<<ceil_div(n, TB), TB, 0, h->stream>>>(n, h->perm.p, h->cen_in.p, h->dir_in.p,
                                                    h->mag_in.p, h->rv_in.p, h->px.p, h->py.p,
                                                    h->pz.p, h->dx.p, h->dy.p, h->dz.p, h->pm.p,
                                                    h->pr.p, nullptr);
  }
  k_fstart<<<ceil_div(h->nfine + 1, TB), TB, 0, h->stream>>>(h->nfine, n, nullptr, h->key_s.p,
                                                             h->fstart.p);
  h->pr_done = false;
  if (n > 0 && h->pr_on) DUST_TRY(redistribute(h));
  k_flags_fine<<<ceil_div(h->nfine, TB), TB, 0, h->stream>>>(G, h->fstart.p, h->prm.min_part,
                                                             h->flags.p);
  for (int l = h->nlev - 2; l >= 0; --l)
    k_flags_level<<<ceil_div(h->lev[l].ncells, TB), TB, 0, h->stream>>>(G, l, h->fstart.p,
                                                                        h->prm.min_part, h->flags.p);
  DUST_TRY(prim_select_flagged_cells(h, 0, h->active.p, h->counts.p));
  DUST_TRY(prim_select_flagged_cells(h, 1, h->leaves.p, h->counts.p + 1));
  DUST_HIP(hipMemcpyAsync(rb->counts, h->counts.p, sizeof(int) * 2, hipMemcpyDeviceToHost,
                          h->stream));
  DUST_HIP(hipMemcpyAsync(&rb->err, h->d_err.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  if (h->pr_done)
    DUST_HIP(hipMemcpyAsync(&rb->ns, h->pr_ns.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  return 0;
}

static int sort_phase_b(FmmHandle *h) {
  const int n = h->n;
  TreeGeom G = make_geom(h);
  const int TB = 256;
  FmmReadback *rb = h->rb.p;
  if (rb->err != rb->init_err) {
    set_error("particle %d resulted outside the octree box", rb->err);
    timer_end(h, DUST_FMM_STAGE_SORT);
    return DUST_E_OUTSIDE;
  }
  h->n_active = rb->counts[0];
  h->n_leaves = rb->counts[1];
  h->ns = h->pr_done ? rb->ns : n;
  k_level_starts<<<1, 64, 0, h->stream>>>(G, h->n_active, h->active.p, h->counts.p + 2);
  DUST_HIP(hipMemcpyAsync(rb->act_start, h->counts.p + 2, sizeof(int) * (h->nlev + 1),
                          hipMemcpyDeviceToHost, h->stream));
  k_fill_int<<<ceil_div(h->ncells_tot, TB), TB, 0, h->stream>>>(h->ncells_tot, h->slot.p, -1);
  if (h->n_active > 0)
    k_slots<<<ceil_div(h->n_active, TB), TB, 0, h->stream>>>(h->n_active, h->active.p, h->slot.p,
                                                                  nullptr);
  // P2P lists
  DUST_TRY(h->p2p_rng.reserve((size_t)std::max(h->n_leaves, 1) * h->p2p_cap));
  DUST_TRY(h->p2p_cnt.reserve(std::max(h->n_leaves, 1)));
  DUST_TRY(h->p2p_tot.reserve(std::max(h->n_leaves, 1)));
  if (h->n_leaves > 0)
    k_p2p_lists<<<ceil_div(h->n_leaves, 4), 256, 0, h->stream>>>(
        G, h->n_leaves, h->leaves.p, h->flags.p, h->fstart.p, h->p2p_cap, h->p2p_rng.p,
        h->p2p_cnt.p, h->p2p_tot.p);
  // evaluation work units and their items (sized after the next synchronisation)
  h->units_cap = h->n_leaves + n / kWave + 1;
  DUST_TRY(h->units.reserve((size_t)h->units_cap * 2));
  DUST_TRY(h->leaf_ros.reserve(std::max(h->n_leaves, 1)));
  DUST_HIP(hipMemsetAsync(h->n_units.p, 0, 2 * sizeof(int), h->stream));
  if (h->n_leaves > 0)
    k_units<<<ceil_div(h->n_leaves, TB), TB, 0, h->stream>>>(
        G, h->n_leaves, h->leaves.p, h->fstart.p, h->p2p_tot.p, h->seg_iters,
        (EvalUnit *)h->units.p, h->n_units.p, h->n_units.p + 1);
  // expansion storage
  size_t na = (size_t)std::max(h->n_active, 1);
  DUST_TRY(h->mp_a.reserve(na * 3 * h->nmon));
  DUST_TRY(h->mp_b.reserve(na * 3 * h->nmon));
  DUST_TRY(h->mp_c.reserve(na * 9 * h->nmon));
  DUST_HIP(hipMemcpyAsync(rb->units, h->n_units.p, 2 * sizeof(int), hipMemcpyDeviceToHost,
                          h->stream));
  return 0;
}

static int sort_phase_c(FmmHandle *h) {
  const int TB = 256;
  const FmmReadback *rb = h->rb.p;
  h->n_unit = rb->units[0];
  h->n_item = rb->units[1];
  std::memcpy(h->act_start, rb->act_start, sizeof(int) * (h->nlev + 1));
  DUST_TRY(h->items.reserve((size_t)std::max(h->n_item, 1) * 4));
  DUST_TRY(h->part.reserve((size_t)std::max(h->n_item, 1) * kWave *
                           ((h->prm.use_vs && h->prm.use_divfilt) ? 9 : 6)));
  if (h->n_unit > 0)
    k_items<<<ceil_div(h->n_unit, TB), TB, 0, h->stream>>>(h->n_unit, (const EvalUnit *)h->units.p,
                                                          h->p2p_tot.p, (EvalItem *)h->items.p, nullptr);
  DUST_HIP(hipGetLastError());
  timer_end(h, DUST_FMM_STAGE_SORT);
  h->sorted = true;
  h->computed = h->applied = false;
  return 0;
}

static int fmm_sort(FmmHandle *h, int *nleaves_out) {
  DUST_TRY(sort_phase_a(h));
  DUST_HIP(hipStreamSynchronize(h->stream));
  DUST_TRY(sort_phase_b(h));
  DUST_HIP(hipStreamSynchronize(h->stream));
  DUST_TRY(sort_phase_c(h));
  if (nleaves_out) *nleaves_out = h->n_leaves;
  return 0;
}

// calculate_multipole (mod_octree.f90:1002-1133)
static int multipole_phase_a(FmmHandle *h) {
  DUST_CHECK(h && h->sorted, DUST_E_ARG, "sort_particles must precede calculate_multipole");
  TreeGeom G = make_geom(h);
  const int nm = h->nmon;
  const int tpb = ((nm + 63) / 64) * 64;
  const int use_vs = h->prm.use_vs;
  if (h->n_active == 0) { h->computed = true; h->applied = false; return 0; }
  // local expansions start from zero (set_leaf/set_branch reset, 1596-1603)
  DUST_HIP(hipMemsetAsync(h->mp_b.p, 0, sizeof(double) * 3 * nm * h->n_active, h->stream));
  if (use_vs)
    DUST_HIP(hipMemsetAsync(h->mp_c.p, 0, sizeof(double) * 9 * nm * h->n_active, h->stream));
  timer_begin(h, DUST_FMM_STAGE_P2M);
  if (h->n_leaves > 0) {
    size_t sh = sizeof(double) * (kP2MChunk * 3 * (h->deg + 1) + kP2MChunk * 4);
    k_p2m<<<h->n_leaves, tpb, sh, h->stream>>>(G, h->deg, nm, h->d_pwr.p, h->leaves.p, h->slot.p,
                                               h->fstart.p, h->px.p, h->py.p, h->pz.p, h->dx.p,
                                               h->dy.p, h->dz.p, h->pm.p, h->mp_a.p);
  }
  timer_end(h, DUST_FMM_STAGE_P2M);
  timer_begin(h, DUST_FMM_STAGE_M2M);
  for (int l = h->nlev - 2; l >= 0; --l) {
    int na = h->act_start[l + 1] - h->act_start[l];
    if (na > 0)
      k_m2m<<<na, tpb, sizeof(double) * 24 * nm, h->stream>>>(G, l, h->deg, nm, h->d_pwr.p,
                                                                h->d_fact.p, h->active.p,
                                       h->act_start[l], h->flags.p, h->slot.p, h->mp_a.p);
  }
  timer_end(h, DUST_FMM_STAGE_M2M);
  if (h->prm.precision == DUST_FMM_TUNED) {
    // translations of every level below the first in one launch (they only
    // read the multipoles); the L2L chain adds the parents afterwards.  The
    // targets are grouped now; the group sizes are read after the next
    // synchronisation (multipole_phase_b).
    timer_begin(h, DUST_FMM_STAGE_M2L);
    DUST_TRY(tuned_pass_prepare(h));
  }
  return 0;
}

static int multipole_phase_b(FmmHandle *h) {
  if (h->n_active == 0) return 0;
  TreeGeom G = make_geom(h);
  const int nm = h->nmon;
  const int tpb = ((nm + 63) / 64) * 64;
  const int use_vs = h->prm.use_vs;
  const bool tuned = h->prm.precision == DUST_FMM_TUNED;
  if (tuned) {
    if (!h->batched) DUST_TRY(tuned_m2l_all(h));  // else in the batch's shared launch
    timer_end(h, DUST_FMM_STAGE_M2L);
  }
  for (int l = 0; l < h->nlev; ++l) {
    int na = h->act_start[l + 1] - h->act_start[l];
    // level 1 has no parent: its interaction list (nbox > 1) is irregular
    // and small, and runs in FP64 in both forms
    if (na > 0 && (!tuned || l == 0)) {
      timer_begin(h, DUST_FMM_STAGE_M2L);
      size_t sh = sizeof(double) * (12 * h->nU + 3 * nm);
      k_m2l_fp64<<<na, tpb, sh, h->stream>>>(G, l, nm, h->nU, use_vs, h->d_pwr.p, h->d_multT.p,
                                             h->d_mnT.p, h->d_kd.p, h->active.p, h->act_start[l],
                                             h->flags.p, h->slot.p, h->mp_a.p, h->mp_b.p,
                                             h->mp_c.p);
      timer_end(h, DUST_FMM_STAGE_M2L);
    }
    if (l + 1 < h->nlev) {
      int nc = h->act_start[l + 2] - h->act_start[l + 1];
      timer_begin(h, DUST_FMM_STAGE_L2L);
      if (nc > 0)
        k_l2l<<<nc, tpb, sizeof(double) * 12 * nm, h->stream>>>(G, l, h->deg, nm, use_vs,
                                                                h->d_pwr.p, h->d_fact.p,
                                         h->active.p, h->act_start[l + 1], h->slot.p, h->mp_b.p,
                                         h->mp_c.p);
      timer_end(h, DUST_FMM_STAGE_L2L);
    }
  }
  h->computed = true;
  h->applied = false;
  return 0;
}

void make_eval_args(const FmmHandle *h, EvalArgs *Ap) {
  EvalArgs &A = *Ap;
  A.G = make_geom(h);
  A.deg = h->deg;
  A.nmon = h->nmon;
  A.use_vs = h->prm.use_vs;
  A.use_vd = h->prm.use_vd;
  A.kernel = h->prm.kernel;
  A.use_tv = h->prm.use_tv;
  A.nu_inf = h->prm.nu_inf;
  for (int d = 0; d < 3; ++d) A.uinf[d] = h->prm.u_inf[d];
  A.pwr = h->d_pwr.p; A.leaves = h->leaves.p; A.slot = h->slot.p; A.fstart = h->fstart.p;
  A.mp_b = h->mp_b.p; A.mp_c = h->mp_c.p;
  A.px = h->px.p; A.py = h->py.p; A.pz = h->pz.p; A.dx = h->dx.p; A.dy = h->dy.p; A.dz = h->dz.p;
  A.pm = h->pm.p; A.pr = h->pr.p; A.rng = h->p2p_rng.p; A.cnt = h->p2p_cnt.p; A.cap = h->p2p_cap;
  A.units = (const EvalUnit *)h->units.p;
  A.items = (const EvalItem *)h->items.p;
  A.n_units = h->n_unit;
  A.n_items = h->n_item;
  A.leaf_ros = h->leaf_ros.p;
  A.near2 = 0.0;
  A.part = h->part.p;
  A.vel = h->vel_s.p; A.str = h->str_s.p; A.tv = h->tv_s.p;
  A.wind = h->wind;
  A.time = h->wind_time;
  A.divf = (h->prm.use_vs && h->prm.use_divfilt) ? 1 : 0;
  A.pstride = A.divf ? 9 : 6;
  A.rot = h->rot_s.p;
}

int eval_leaf_ros(FmmHandle *h, const EvalArgs &A) {
  const int nl = h->gr.active ? h->gr.leaf_cap : h->n_leaves;  // graph form: by capacity
  if (!(h->prm.use_vd && h->prm.use_tv) || nl == 0) return 0;
  size_t sh = sizeof(double) * kEvalWaves * 12 * h->nmon + sizeof(int) * 3 * h->nmon;
  k_leaf_ros<<<ceil_div(nl, kEvalWaves), kWave * kEvalWaves, sh, h->stream>>>(A, nl);
  DUST_HIP(hipGetLastError());
  return 0;
}

int eval_fp64_launch(FmmHandle *h, const EvalArgs &A) {
  DUST_TRY(eval_leaf_ros(h, A));
  if (A.n_items > 0)
    k_p2p_fp64<<<ceil_div(A.n_items, kEvalWaves), kWave * kEvalWaves, 0, h->stream>>>(A);
  DUST_HIP(hipGetLastError());
  return 0;
}

int eval_finish_launch(FmmHandle *h, const EvalArgs &A) {
  size_t sh = eval_finish_lds(h->nmon);
  if (A.n_units > 0)
    k_eval_finish<<<ceil_div(A.n_units, kEvalWaves), kWave * kEvalWaves, sh, h->stream>>>(A);
  DUST_HIP(hipGetLastError());
  return 0;
}

static int fmm_multipole(FmmHandle *h) {
  DUST_TRY(multipole_phase_a(h));
  if (h->prm.precision == DUST_FMM_TUNED && h->n_active > 0)
    DUST_HIP(hipStreamSynchronize(h->stream));  // target group sizes
  return multipole_phase_b(h);
}

// apply_multipole (mod_octree.f90:1139-1429), particle terms + free stream
static int fmm_apply(FmmHandle *h) {
  DUST_CHECK(h && h->computed, DUST_E_ARG, "calculate_multipole must precede apply_multipole");
  const int n = h->n;
  h->applied = true;
  if (n == 0) return 0;
  if (h->ns < n)  // freed by the redistribution: not evaluated
    k_pr_zero<<<ceil_div(n, 256), 256, 0, h->stream>>>(n, h->pr_keep.p, h->vel.p, h->str.p,
                                                       h->tv.p, h->rot.p);
  if (h->ns == 0) return 0;
  EvalArgs A;
  make_eval_args(h, &A);
  timer_begin(h, DUST_FMM_STAGE_P2P);
  if (h->batched) {
    // the P2P items ran in the batch's shared launch
  } else if (h->prm.precision == DUST_FMM_TUNED) {
    DUST_TRY(tuned_eval(h, A));
  } else {
    DUST_TRY(eval_fp64_launch(h, A));
  }
  timer_end(h, DUST_FMM_STAGE_P2P);
  // L2P, the sum of the items' partial sums and the free stream
  timer_begin(h, DUST_FMM_STAGE_L2P);
  DUST_TRY(eval_finish_launch(h, A));
  timer_end(h, DUST_FMM_STAGE_L2P);
  k_scatter<<<ceil_div(h->ns, 256), 256, 0, h->stream>>>(h->ns, h->perm.p, h->vel_s.p, h->str_s.p,
                                                         h->tv_s.p, h->vel.p, h->str.p, h->tv.p,
                                                         nullptr);
  if (A.divf)
    k_scatter3<<<ceil_div(h->ns, 256), 256, 0, h->stream>>>(h->ns, h->perm.p, h->rot_s.p,
                                                            h->rot.p, nullptr);
  return 0;
}


// ---------------------------------------------------------------------------
// apply_multipole_panels (mod_octree.f90:1434-1510) with the panel binning of
// sort_particles (777-820): the particle velocity at panel centres from the
// tree of the last pass, added into uvort.
// ---------------------------------------------------------------------------

// Leaf of every centre: its finest cell (ceiling, as for the particles), then
// the first ancestor-or-self flagged leaf (every point of the box lies under
// one: a level-1 cell with no active child is a leaf; -1 is kept for an empty
// tree).
__global__ void k_pan_leaf(TreeGeom G, int m, const double *__restrict__ cen,
                           const unsigned char *__restrict__ flags,
                           const int *__restrict__ leaves, int n_leaves, int *pleaf, int *err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int L = G.nlev - 1;
  int id[3];
  bool bad = false;
  for (int d = 0; d < 3; ++d) {
    double q = ceil((cen[3 * i + d] - G.xmin[d]) / G.csize_fine);
    if (!(q >= 1.0 && q <= (double)G.ncl[L][d])) bad = true;
    id[d] = bad ? 1 : (int)q - 1;
  }
  pleaf[i] = -1;
  if (bad) {
    atomicMin(err, i);
    return;
  }
  for (int l = L; l >= 0; --l) {
    int sh = L - l;
    long g = G.off[l] + cell_encode(G, l, id[0] >> sh, id[1] >> sh, id[2] >> sh);
    if (flags[g] == kLeaf) {
      int lo = 0, hi = n_leaves - 1;  // leaves are in ascending cell order
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (leaves[mid] < g) lo = mid + 1;
        else hi = mid;
      }
      if (n_leaves > 0 && leaves[lo] == g) pleaf[i] = lo;
      return;
    }
  }
}

// One wavefront per panel: the lanes split the leaf's P2P list (velocity of
// compute_vel_vortpart with the handle's particle core), a fixed-order
// reduction, then lane 0 adds the L2P velocity and writes uvort += .
__global__ void __launch_bounds__(256)
    k_pan_eval(EvalArgs A, int m, const double *__restrict__ cen, const int *__restrict__ pleaf,
               double *uvort) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= m) return;
  const int lv = pleaf[i];
  if (lv < 0) return;
  const double x = cen[3 * i], y = cen[3 * i + 1], z = cen[3 * i + 2];
  const int2 *rl = &A.rng[(long)lv * A.cap];
  const int nr = A.cnt[lv];
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  for (int r = 0; r < nr; ++r) {
    const int2 rg = rl[r];
    for (int q = rg.x + lane; q < rg.y; q += kWave) {
      double d0 = x - A.px[q], d1 = y - A.py[q], d2 = z - A.pz[q];
      double rr = A.pr[q];
      double c;
      if (A.kernel == DUST_KERNEL_GAUSSIAN) {
        double cg, dg;
        gauss_cd(d0 * d0 + d1 * d1 + d2 * d2, rr, cg, dg);
        c = A.pm[q] * cg;
      } else {
        double R2 = d0 * d0 + d1 * d1 + d2 * d2 + rr * rr;
        c = -A.pm[q] / (R2 * sqrt(R2));
      }
      double a0 = A.dx[q], a1 = A.dy[q], a2 = A.dz[q];
      v0 += c * (d1 * a2 - d2 * a1);
      v1 += c * (d2 * a0 - d0 * a2);
      v2 += c * (d0 * a1 - d1 * a0);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    v0 += __shfl_xor(v0, o, kWave);
    v1 += __shfl_xor(v1, o, kWave);
    v2 += __shfl_xor(v2, o, kWave);
  }
  if (lane != 0) return;
  const long g = A.leaves[lv];
  const int lvl = lvl_of(A.G, g);
  int ix, iy, iz;
  cell_decode(A.G, lvl, g - A.G.off[lvl], ix, iy, iz);
  const double cx = cell_center(A.G, lvl, 0, ix), cy = cell_center(A.G, lvl, 1, iy),
               cz = cell_center(A.G, lvl, 2, iz);
  const double *sb = A.mp_b + (long)A.slot[g] * 3 * A.nmon;
  double v[3], gr[9];
  l2p_point(A.deg, A.nmon, 0, sb, sb, A.pwr, x - cx, y - cy, z - cz, v, gr);
  const double f = 1.0 / (4.0 * kPi);
  uvort[3 * i] += v[0] + v0 * f;
  uvort[3 * i + 1] += v[1] + v1 * f;
  uvort[3 * i + 2] += v[2] + v2 * f;
}

// leaf of every panel centre and the out-of-box report of apply_multipole_panels
static DevBuf<int> g_pan_leaf, g_pan_err;

void fmm_release() {
  g_pan_leaf.release();
  g_pan_err.release();
  tuned_batch_release();
}

static int fmm_apply_panels(FmmHandle *h, int m, const double *cen, double *uvort) {
  DUST_CHECK(h && h->computed, DUST_E_ARG,
             "calculate_multipole must precede apply_multipole_panels");
  DUST_CHECK(m >= 0 && (m == 0 || (cen && uvort)), DUST_E_ARG, "invalid panel arguments");
  if (m == 0) return 0;
  DevBuf<int> &pleaf = g_pan_leaf, &perr = g_pan_err;
  DUST_TRY(pleaf.reserve(m));
  DUST_TRY(perr.reserve(1));
  DUST_TRY(h->rb.reserve(1));
  TreeGeom G = make_geom(h);
  k_fill_int<<<1, 32, 0, h->stream>>>(1, perr.p, 0x7fffffff);
  k_pan_leaf<<<ceil_div(m, 256), 256, 0, h->stream>>>(G, m, cen, h->flags.p, h->leaves.p,
                                                     h->n_leaves, pleaf.p, perr.p);
  DUST_HIP(hipMemcpyAsync(&h->rb.p->err, perr.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  DUST_HIP(hipStreamSynchronize(h->stream));
  DUST_CHECK(h->rb.p->err == 0x7fffffff, DUST_E_OUTSIDE,
             "panel centre %d outside the octree box", h->rb.p->err);
  if (h->ns == 0 || h->n_leaves == 0) return 0;
  EvalArgs A;
  make_eval_args(h, &A);
  k_pan_eval<<<ceil_div(m, 4), 256, 0, h->stream>>>(A, m, cen, pleaf.p, uvort);
  DUST_HIP(hipGetLastError());
  return 0;
}
// ---------------------------------------------------------------------------
// Graph form of the pass: sort_particles, calculate_multipole and
// apply_multipole (mod_octree.f90:703-1429) enqueued with every count on the
// device.  The grids and buffers are sized by the capacities of
// dust_fmm_graph_reserve; each kernel reads its count (particles, active
// cells, leaves, level starts, units, items, M2L queue length) from
// GraphState::dc and does nothing past it, or at all once the stop word is
// set (a capacity exceeded), which the status call reports after the run.
// Same kernels, same summation order as the host-ordered pass.
// ---------------------------------------------------------------------------
__global__ void k_graph_begin(const int *__restrict__ n_in, int cap, int *dc) {
  const int t = threadIdx.x;
  if (t >= kGcSize) return;
  const int n = *n_in;
  const bool bad = n < 0 || n > cap;
  int v = 0;
  if (t == kGcN) v = bad ? 0 : n;
  else if (t == kGcOvf) v = bad ? 1 : 0;
  else if (t == kGcErr) v = 0x7fffffff;
  dc[t] = v;
}

struct LevCaps {
  int c[kMaxLevels];
};

// The lists against the capacities the grids and buffers were sized for.
__global__ void k_graph_check(int nlev, LevCaps lc, int act_cap, int leaf_cap, int *dc) {
  if (threadIdx.x != 0) return;
  int bits = 0;
  if (dc[kGcActive] > act_cap || dc[kGcLeaves] > leaf_cap) bits |= 2;
  for (int l = 0; l < nlev; ++l)
    if (dc[kGcActStart + l + 1] - dc[kGcActStart + l] > lc.c[l]) bits |= 4;
  if (bits) atomicOr(&dc[kGcOvf], bits);
}

// Zero the first dc[idx] * per entries (the local expansions of the active
// cells).
__global__ void k_zero_dc(double *p, long per, int idx, const int *__restrict__ dc) {
  const long n = dc[kGcOvf] ? 0 : (long)dc[idx] * per;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

struct GraphActive {
  FmmHandle *h;
  explicit GraphActive(FmmHandle *hh) : h(hh) { h->gr.active = true; }
  ~GraphActive() { h->gr.active = false; }
};

static std::vector<const void *> graph_buffers(const FmmHandle *h) {
  std::vector<const void *> v = {
      h->key.p, h->key_s.p, h->idx_in.p, h->perm.p, h->px.p, h->py.p, h->pz.p, h->dx.p,
      h->dy.p, h->dz.p, h->pm.p, h->pr.p, h->fstart.p, h->flags.p, h->slot.p, h->active.p,
      h->leaves.p, h->p2p_rng.p, h->p2p_cnt.p, h->p2p_tot.p, h->units.p, h->items.p, h->part.p,
      h->leaf_ros.p, h->mp_a.p, h->mp_b.p, h->mp_c.p, h->vel_s.p, h->str_s.p, h->tv_s.p,
      h->rot_s.p, h->rot.p, h->vel.p, h->str.p, h->tv.p, h->d_pwr.p, h->d_multT.p, h->d_mnT.p,
      h->d_kd.p, h->d_fact.p, h->gr.dc.p, h->gr.rs_hist.p, h->gr.sel_cnt.p};
  if (h->prm.precision == DUST_FMM_TUNED) tuned_graph_buffers(h, v);
  return v;
}

static int graph_reserve(FmmHandle *h, double slack) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  DUST_CHECK(h->sorted && h->computed, DUST_E_ARG,
             "graph reserve sizes its lists from a completed pass on the handle");
  DUST_CHECK(slack >= 1.0 && slack <= 64.0, DUST_E_ARG, "slack %g outside [1, 64]", slack);
  DUST_CHECK(!h->pr_on, DUST_E_ARG,
             "graph form: the particle redistribution is not available (set_redistribution off)");
  DUST_CHECK(h->wind.type == DUST_GUST_NONE, DUST_E_ARG,
             "graph form: a gust's time would be fixed at capture (uniform free stream only)");
  GraphState &R = h->gr;
  if (R.exec) (void)hipGraphExecDestroy(R.exec);
  if (R.graph) (void)hipGraphDestroy(R.graph);
  R.exec = nullptr;
  R.graph = nullptr;
  R.reserved = false;
  R.fp.clear();
  auto grow = [&](long v, long hi) {
    const long c = (long)std::ceil((double)v * slack) + 64;
    return (int)std::max(0L, std::min(c, hi));
  };
  R.n_cap = h->cap;
  R.act_cap = grow(h->n_active, h->ncells_tot);
  R.leaf_cap = grow(h->n_leaves, h->ncells_tot);
  // a level may hold any share of the active cells: its grids are sized by
  // the smaller of its cell count and the active-cell capacity (the blocks
  // past the level's count return at once)
  for (int l = 0; l < kMaxLevels; ++l)
    R.lev_cap[l] = l < h->nlev ? (int)std::min<long>(h->lev[l].ncells, R.act_cap) : 0;
  R.unit_cap = R.leaf_cap + R.n_cap / kWave + 1;
  R.item_cap = grow(h->n_item, (long)INT32_MAX / 2);
  const int pstride = (h->prm.use_vs && h->prm.use_divfilt) ? 9 : 6;
  const size_t na = std::max(R.act_cap, 1), nl = std::max(R.leaf_cap, 1);
  DUST_TRY(R.dc.reserve(kGcSize));
  DUST_TRY(R.rs_hist.reserve(gprim_hist_ints(std::max(R.n_cap, R.act_cap))));
  DUST_TRY(R.sel_cnt.reserve(gprim_sel_ints(h->ncells_tot)));
  DUST_TRY(h->active.reserve(na));
  DUST_TRY(h->leaves.reserve(nl));
  DUST_TRY(h->p2p_rng.reserve(nl * h->p2p_cap));
  DUST_TRY(h->p2p_cnt.reserve(nl));
  DUST_TRY(h->p2p_tot.reserve(nl));
  DUST_TRY(h->leaf_ros.reserve(nl));
  DUST_TRY(h->units.reserve((size_t)R.unit_cap * 2));
  DUST_TRY(h->items.reserve((size_t)std::max(R.item_cap, 1)));
  DUST_TRY(h->part.reserve((size_t)std::max(R.item_cap, 1) * kWave * pstride));
  DUST_TRY(h->mp_a.reserve(na * 3 * h->nmon));
  DUST_TRY(h->mp_b.reserve(na * 3 * h->nmon));
  DUST_TRY(h->mp_c.reserve(na * 9 * h->nmon));
  if (h->prm.precision == DUST_FMM_TUNED) DUST_TRY(tuned_graph_reserve(h));
  R.fp = graph_buffers(h);
  R.reserved = true;
  return 0;
}

static int graph_enqueue(FmmHandle *h, const int *d_n, const double *cen, const double *dir,
                         const double *mag, const double *rv, double *vel, double *str) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  GraphState &R = h->gr;
  DUST_CHECK(R.reserved, DUST_E_ARG, "dust_fmm_graph_reserve must precede the graph form");
  DUST_CHECK(d_n && cen && dir && mag && rv && vel, DUST_E_ARG, "null device array");
  DUST_CHECK(graph_buffers(h) == R.fp, DUST_E_ARG,
             "handle buffers were reallocated since dust_fmm_graph_reserve");
  GraphActive on(h);
  h->sorted = h->computed = h->applied = false;  // host-side counts are not this pass's
  const TreeGeom G = make_geom(h);
  int *dc = R.dc.p;
  const int TB = 256, cap = std::max(R.n_cap, 1);
  const int nm = h->nmon, tpb = ((nm + 63) / 64) * 64;
  const bool tuned = h->prm.precision == DUST_FMM_TUNED;
  // sort_particles
  k_graph_begin<<<1, 64, 0, h->stream>>>(d_n, R.n_cap, dc);
  k_keys<<<ceil_div(cap, TB), TB, 0, h->stream>>>(G, 0, cen, h->key.p, h->idx_in.p, dc + kGcErr);
  int bits = 0;
  while ((1L << bits) < h->nfine) ++bits;
  if (bits == 0) bits = 1;
  DCount cn{dc + kGcN, nullptr, nullptr, R.n_cap};
  DUST_TRY(gprim_sort_pairs_u32(h, h->key.p, h->idx_in.p, h->key_s.p, h->perm.p, cn, bits));
  k_gather<<<ceil_div(cap, TB), TB, 0, h->stream>>>(0, h->perm.p, cen, dir, mag, rv, h->px.p,
                                                    h->py.p, h->pz.p, h->dx.p, h->dy.p, h->dz.p,
                                                    h->pm.p, h->pr.p, dc + kGcN);
  k_fstart<<<ceil_div(h->nfine + 1, TB), TB, 0, h->stream>>>(h->nfine, 0, dc + kGcN, h->key_s.p,
                                                             h->fstart.p);
  k_flags_fine<<<ceil_div(h->nfine, TB), TB, 0, h->stream>>>(G, h->fstart.p, h->prm.min_part,
                                                             h->flags.p);
  for (int l = h->nlev - 2; l >= 0; --l)
    k_flags_level<<<ceil_div(h->lev[l].ncells, TB), TB, 0, h->stream>>>(G, l, h->fstart.p,
                                                                        h->prm.min_part, h->flags.p);
  DUST_TRY(gprim_select_flagged_cells(h, 0, h->active.p, R.act_cap, dc + kGcActive));
  DUST_TRY(gprim_select_flagged_cells(h, 1, h->leaves.p, R.leaf_cap, dc + kGcLeaves));
  k_level_starts<<<1, 64, 0, h->stream>>>(G, 0, h->active.p, dc + kGcActStart);
  LevCaps lc;
  for (int l = 0; l < kMaxLevels; ++l) lc.c[l] = R.lev_cap[l];
  k_graph_check<<<1, 64, 0, h->stream>>>(h->nlev, lc, R.act_cap, R.leaf_cap, dc);
  k_fill_int<<<ceil_div(h->ncells_tot, TB), TB, 0, h->stream>>>(h->ncells_tot, h->slot.p, -1);
  if (R.act_cap > 0)
    k_slots<<<ceil_div(R.act_cap, TB), TB, 0, h->stream>>>(0, h->active.p, h->slot.p, dc);
  if (R.leaf_cap > 0) {
    k_p2p_lists<<<ceil_div(R.leaf_cap, 4), 256, 0, h->stream>>>(
        G, 0, h->leaves.p, h->flags.p, h->fstart.p, h->p2p_cap, h->p2p_rng.p, h->p2p_cnt.p,
        h->p2p_tot.p);
    k_units<<<ceil_div(R.leaf_cap, TB), TB, 0, h->stream>>>(
        G, 0, h->leaves.p, h->fstart.p, h->p2p_tot.p, h->seg_iters, (EvalUnit *)h->units.p,
        dc + kGcUnits, dc + kGcItems);
  }
  k_items<<<ceil_div(R.unit_cap, TB), TB, 0, h->stream>>>(0, (const EvalUnit *)h->units.p,
                                                          h->p2p_tot.p, (EvalItem *)h->items.p, dc);
  // calculate_multipole
  k_zero_dc<<<2048, TB, 0, h->stream>>>(h->mp_b.p, 3L * nm, kGcActive, dc);
  if (h->prm.use_vs) k_zero_dc<<<2048, TB, 0, h->stream>>>(h->mp_c.p, 9L * nm, kGcActive, dc);
  if (R.leaf_cap > 0) {
    size_t sh = sizeof(double) * (kP2MChunk * 3 * (h->deg + 1) + kP2MChunk * 4);
    k_p2m<<<R.leaf_cap, tpb, sh, h->stream>>>(G, h->deg, nm, h->d_pwr.p, h->leaves.p, h->slot.p,
                                              h->fstart.p, h->px.p, h->py.p, h->pz.p, h->dx.p,
                                              h->dy.p, h->dz.p, h->pm.p, h->mp_a.p);
  }
  for (int l = h->nlev - 2; l >= 0; --l)
    if (R.lev_cap[l] > 0)
      k_m2m<<<R.lev_cap[l], tpb, sizeof(double) * 24 * nm, h->stream>>>(
          G, l, h->deg, nm, h->d_pwr.p, h->d_fact.p, h->active.p, 0, h->flags.p, h->slot.p,
          h->mp_a.p);
  if (tuned) DUST_TRY(tuned_graph_m2l(h));
  for (int l = 0; l < h->nlev; ++l) {
    if ((!tuned || l == 0) && R.lev_cap[l] > 0) {
      size_t sh = sizeof(double) * (12 * h->nU + 3 * nm);
      k_m2l_fp64<<<R.lev_cap[l], tpb, sh, h->stream>>>(
          G, l, nm, h->nU, h->prm.use_vs, h->d_pwr.p, h->d_multT.p, h->d_mnT.p, h->d_kd.p,
          h->active.p, 0, h->flags.p, h->slot.p, h->mp_a.p, h->mp_b.p, h->mp_c.p);
    }
    if (l + 1 < h->nlev && R.lev_cap[l + 1] > 0)
      k_l2l<<<R.lev_cap[l + 1], tpb, sizeof(double) * 12 * nm, h->stream>>>(
          G, l, h->deg, nm, h->prm.use_vs, h->d_pwr.p, h->d_fact.p, h->active.p, 0, h->slot.p,
          h->mp_b.p, h->mp_c.p);
  }
  // apply_multipole: grids by capacity, counts from dc
  EvalArgs A;
  make_eval_args(h, &A);
  A.n_units = R.unit_cap;
  A.n_items = R.item_cap;
  if (tuned) DUST_TRY(tuned_eval(h, A));
  else DUST_TRY(eval_fp64_launch(h, A));
  DUST_TRY(eval_finish_launch(h, A));
  k_scatter<<<ceil_div(cap, 256), 256, 0, h->stream>>>(0, h->perm.p, h->vel_s.p, h->str_s.p,
                                                       h->tv_s.p, vel, str ? str : h->str.p,
                                                       h->tv.p, dc + kGcN);
  if (A.divf)
    k_scatter3<<<ceil_div(cap, 256), 256, 0, h->stream>>>(0, h->perm.p, h->rot_s.p, h->rot.p,
                                                          dc + kGcN);
  DUST_HIP(hipGetLastError());
  return 0;
}

static void graph_drop(GraphState &R) {
  if (R.exec) (void)hipGraphExecDestroy(R.exec);
  if (R.graph) (void)hipGraphDestroy(R.graph);
  R.exec = nullptr;
  R.graph = nullptr;
}

static int graph_capture(FmmHandle *h, const int *d_n, const double *cen, const double *dir,
                         const double *mag, const double *rv, double *vel, double *str) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  GraphState &R = h->gr;
  graph_drop(R);
  DUST_HIP(hipStreamSynchronize(h->stream));
  DUST_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const int rc = graph_enqueue(h, d_n, cen, dir, mag, rv, vel, str);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    set_error("stream capture of the pass failed: %s", hipGetErrorString(e));
    return DUST_E_HIP;
  }
  R.graph = g;
  const hipError_t ei = hipGraphInstantiate(&R.exec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    graph_drop(R);
    set_error("graph instantiation failed: %s", hipGetErrorString(ei));
    return DUST_E_HIP;
  }
  return 0;
}

static int graph_launch(FmmHandle *h) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  GraphState &R = h->gr;
  DUST_CHECK(R.exec, DUST_E_ARG, "no captured pass on the handle");
  DUST_CHECK(graph_buffers(h) == R.fp, DUST_E_ARG,
             "handle buffers were reallocated since the capture");
  h->sorted = h->computed = h->applied = false;
  DUST_HIP(hipGraphLaunch(R.exec, h->stream));
  return 0;
}

static int graph_status(FmmHandle *h, int *counts) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  GraphState &R = h->gr;
  DUST_CHECK(R.reserved, DUST_E_ARG, "no graph form reserved on the handle");
  int c[kGcSize];
  DUST_HIP(hipMemcpyAsync(c, R.dc.p, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  DUST_HIP(hipStreamSynchronize(h->stream));
  if (counts) {
    counts[0] = c[kGcN];
    counts[1] = c[kGcActive];
    counts[2] = c[kGcLeaves];
    counts[3] = c[kGcItems];
  }
  if (c[kGcOvf] & 1) {
    set_error("graph pass: particle count outside 0..%d", R.n_cap);
    return DUST_E_CAPACITY;
  }
  if (c[kGcErr] != 0x7fffffff) {
    set_error("particle %d resulted outside the octree box", c[kGcErr]);
    return DUST_E_OUTSIDE;
  }
  if (c[kGcOvf]) {
    set_error("graph pass: capacity exceeded (%s%s%s%s; %d active cells / %d, %d leaves / %d, "
              "%d items / %d): reserve again with more slack and capture again",
              (c[kGcOvf] & 2) ? "cells " : "", (c[kGcOvf] & 4) ? "level " : "",
              (c[kGcOvf] & 8) ? "items " : "", (c[kGcOvf] & 16) ? "M2L jobs" : "", c[kGcActive],
              R.act_cap, c[kGcLeaves], R.leaf_cap, c[kGcItems], R.item_cap);
    return DUST_E_CAPACITY;
  }
  return 0;
}
}  // namespace dust

// ===========================================================================
// C ABI
// ===========================================================================
using namespace dust;

struct dust_fmm_s {
  FmmHandle *h;
};

extern "C" {

const char *dust_last_error(void) { return g_err; }

const char *dust_version(void) { return "libdust_gpu 0.2 (gfx950)"; }

int dust_shutdown(void) {
  // the workspaces may still be in use by queued work
  DUST_HIP(hipDeviceSynchronize());
  aic_release();
  asm_release();
  elem_release();
  fmm_release();
  lu_release();
  mixed_release();
  pres_release();
  wake_release();
  return 0;
}

int dust_fmm_create(const dust_fmm_params *prm, int capacity, dust_fmm_t *out) {
  DUST_CHECK(out, DUST_E_ARG, "null output handle");
  FmmHandle *h = nullptr;
  int rc = fmm_create(prm, capacity, &h);
  if (rc) return rc;
  *out = new dust_fmm_s{h};
  return 0;
}

int dust_fmm_destroy(dust_fmm_t h) {
  if (!h) return 0;
  fmm_destroy(h->h);
  delete h;
  return 0;
}

int dust_fmm_set_particles(dust_fmm_t h, int n, const double *cen, const double *dir,
                           const double *mag, const double *rv) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return fmm_set_particles(h->h, n, cen, dir, mag, rv, hipMemcpyHostToDevice);
}

int dust_fmm_set_particles_dev(dust_fmm_t h, int n, const double *cen, const double *dir,
                               const double *mag, const double *rv) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return fmm_set_particles(h->h, n, cen, dir, mag, rv, hipMemcpyDeviceToDevice);
}

int dust_fmm_sort_particles(dust_fmm_t h, int *nleaves) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return fmm_sort(h->h, nleaves);
}

int dust_fmm_set_redistribution(dust_fmm_t hh, int on, int lvl_solid, double ratio, int m,
                                const double *elem_cen) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  if (!on) {
    h->pr_on = 0;
    h->pr_m = 0;
    return 0;
  }
  DUST_CHECK(lvl_solid >= 1 && lvl_solid <= h->nlev, DUST_E_ARG,
             "octree_level_solid %d outside 1..%d", lvl_solid, h->nlev);
  DUST_CHECK(ratio > 0.0, DUST_E_ARG, "particles_redistribution_ratio must be positive");
  DUST_CHECK(m >= 0 && (m == 0 || elem_cen), DUST_E_ARG, "missing element centres");
  DUST_TRY(h->pr_cen.reserve((size_t)3 * std::max(m, 1)));
  if (m > 0) {
    DUST_HIP(hipMemcpyAsync(h->pr_cen.p, elem_cen, sizeof(double) * 3 * m, hipMemcpyDefault,
                            h->stream));
    DUST_HIP(hipStreamSynchronize(h->stream));
  }
  h->pr_on = 1;
  h->pr_lvl = lvl_solid - 1;
  h->pr_ratio = ratio;
  h->pr_m = m;
  return 0;
}

static int redistributed_copy(FmmHandle *h, unsigned char *keep, double *dir, double *mag,
                              hipMemcpyKind kind) {
  DUST_CHECK(h->sorted, DUST_E_ARG, "sort_particles must be called first");
  const int n = h->n;
  if (n == 0) return 0;
  if (h->pr_done) {
    if (keep) DUST_HIP(hipMemcpyAsync(keep, h->pr_keep.p, n, kind, h->stream));
    if (dir) DUST_HIP(hipMemcpyAsync(dir, h->pr_dir.p, sizeof(double) * 3 * n, kind, h->stream));
    if (mag) DUST_HIP(hipMemcpyAsync(mag, h->pr_mag.p, sizeof(double) * n, kind, h->stream));
  } else {
    // no redistribution in the last sort: every particle kept, as given
    if (keep) {
      if (kind == hipMemcpyDeviceToHost) std::memset(keep, 1, n);
      else k_fill_u8<<<ceil_div(n, 256), 256, 0, h->stream>>>(n, keep, 1);
    }
    if (dir) DUST_HIP(hipMemcpyAsync(dir, h->dir_in.p, sizeof(double) * 3 * n, kind, h->stream));
    if (mag) DUST_HIP(hipMemcpyAsync(mag, h->mag_in.p, sizeof(double) * n, kind, h->stream));
  }
  return 0;
}

int dust_fmm_redistributed(dust_fmm_t hh, int *n_kept, unsigned char *keep, double *dir,
                           double *mag) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  DUST_TRY(redistributed_copy(h, keep, dir, mag, hipMemcpyDeviceToHost));
  DUST_HIP(hipStreamSynchronize(h->stream));
  if (n_kept) *n_kept = h->ns;
  return 0;
}

int dust_fmm_redistributed_dev(dust_fmm_t hh, unsigned char *keep, double *dir, double *mag) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  return redistributed_copy(hh->h, keep, dir, mag, hipMemcpyDeviceToDevice);
}

int dust_fmm_calculate_multipole(dust_fmm_t h) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  DUST_TRY(fmm_multipole(h->h));
  return 0;
}

int dust_fmm_apply_multipole_dev(dust_fmm_t h) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return fmm_apply(h->h);
}

int dust_fmm_apply_multipole(dust_fmm_t hh, double *vel, double *stretch, double *turbvisc) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  DUST_TRY(fmm_apply(h));
  int n = h->n;
  if (n > 0) {
    if (vel) DUST_HIP(hipMemcpyAsync(vel, h->vel.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
    if (stretch)
      DUST_HIP(hipMemcpyAsync(stretch, h->str.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
    if (turbvisc)
      DUST_HIP(hipMemcpyAsync(turbvisc, h->tv.p, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  }
  DUST_HIP(hipStreamSynchronize(h->stream));
  timer_collect(h);
  return 0;
}

static int rotu_copy(FmmHandle *h, double *rotu, hipMemcpyKind kind) {
  DUST_CHECK(h->prm.use_vs && h->prm.use_divfilt, DUST_E_ARG,
             "rotu needs a handle created with use_vs and use_divfilt");
  DUST_CHECK(h->applied, DUST_E_ARG, "apply_multipole must precede rotu");
  if (h->n > 0 && rotu)
    DUST_HIP(hipMemcpyAsync(rotu, h->rot.p, sizeof(double) * 3 * h->n, kind, h->stream));
  return 0;
}

int dust_fmm_rotu(dust_fmm_t hh, double *rotu) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  DUST_TRY(rotu_copy(hh->h, rotu, hipMemcpyDeviceToHost));
  DUST_HIP(hipStreamSynchronize(hh->h->stream));
  return 0;
}

int dust_fmm_rotu_dev(dust_fmm_t hh, double *rotu) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  return rotu_copy(hh->h, rotu, hipMemcpyDeviceToDevice);
}

int dust_fmm_set_wind(dust_fmm_t hh, const dust_wind_params *w, double time) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  if (!w) {
    h->wind = WindDev{};
    for (int e = 0; e < 3; ++e) h->wind.u_inf[e] = h->prm.u_inf[e];
    h->wind.type = DUST_GUST_NONE;
    h->wind_time = 0.0;
    return 0;
  }
  DUST_CHECK(w->gust_type >= DUST_GUST_NONE && w->gust_type <= DUST_GUST_LINEAR, DUST_E_ARG,
             "unknown gust type %d", w->gust_type);
  h->wind = wind_dev(*w);
  h->wind_time = time;
  return 0;
}

int dust_fmm_results_dev(dust_fmm_t hh, double *vel, double *stretch, double *turbvisc) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  int n = h->n;
  if (n == 0) return 0;
  if (vel) DUST_HIP(hipMemcpyAsync(vel, h->vel.p, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, h->stream));
  if (stretch)
    DUST_HIP(hipMemcpyAsync(stretch, h->str.p, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, h->stream));
  if (turbvisc)
    DUST_HIP(hipMemcpyAsync(turbvisc, h->tv.p, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
  return 0;
}

int dust_fmm_apply_multipole_panels_dev(dust_fmm_t hh, int m, const double *cen,
                                        double *uvort) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  return fmm_apply_panels(hh->h, m, cen, uvort);
}

int dust_fmm_apply_multipole_panels(dust_fmm_t hh, int m, const double *cen, double *uvort) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  DUST_CHECK(m >= 0 && (m == 0 || (cen && uvort)), DUST_E_ARG, "invalid panel arguments");
  if (m == 0) return 0;
  double *d = nullptr;
  DUST_HIP(hipMalloc(&d, sizeof(double) * 6 * (size_t)m));
  int rc = 0;
  if (hipMemcpy(d, cen, sizeof(double) * 3 * m, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + 3 * (size_t)m, uvort, sizeof(double) * 3 * m, hipMemcpyHostToDevice) !=
          hipSuccess) {
    set_error("copy of the panel centres failed");
    rc = DUST_E_HIP;
  }
  if (rc == 0) rc = fmm_apply_panels(h, m, d, d + 3 * (size_t)m);
  if (rc == 0 &&
      hipMemcpyAsync(uvort, d + 3 * (size_t)m, sizeof(double) * 3 * m, hipMemcpyDeviceToHost,
                     h->stream) != hipSuccess) {
    set_error("copy of uvort failed");
    rc = DUST_E_HIP;
  }
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(d);
  return rc;
}

int dust_fmm_pass_dev(dust_fmm_t hh, int *nleaves) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  DUST_TRY(fmm_sort(hh->h, nleaves));
  DUST_TRY(fmm_multipole(hh->h));
  DUST_TRY(fmm_apply(hh->h));
  return 0;
}

// Several independent cases: the passes advance in lockstep, phase by phase
// (sort a / b / c, multipole a / b, apply), each handle on its own stream; one
// synchronisation of every stream separates two phases, so the kernels of
// different cases overlap and the host round trips are shared.  The tuned
// passes (FP32 P2P, same use_vs and core) share two persistent launches: the
// M2L jobs of all of them and the P2P items of all of them, on the first such
// handle's stream; every handle's L2L chain waits for the M2L launch and its
// L2P / sums for the P2P launch.
int dust_fmm_pass_batch_dev(int nh, dust_fmm_t *hs, int *nleaves) {
  DUST_CHECK(nh >= 0 && (nh == 0 || hs), DUST_E_ARG, "invalid batch arguments");
  for (int i = 0; i < nh; ++i) DUST_CHECK(hs[i], DUST_E_ARG, "null handle %d in batch", i);
  auto sync_all = [&]() -> int {
    for (int i = 0; i < nh; ++i) DUST_HIP(hipStreamSynchronize(hs[i]->h->stream));
    return 0;
  };
  for (int i = 0; i < nh; ++i) DUST_TRY(sort_phase_a(hs[i]->h));
  DUST_TRY(sync_all());
  for (int i = 0; i < nh; ++i) DUST_TRY(sort_phase_b(hs[i]->h));
  DUST_TRY(sync_all());
  for (int i = 0; i < nh; ++i) {
    DUST_TRY(sort_phase_c(hs[i]->h));
    DUST_TRY(multipole_phase_a(hs[i]->h));
    if (nleaves) nleaves[i] = hs[i]->h->n_leaves;
  }
  DUST_TRY(sync_all());
  // the passes of the shared launches
  std::vector<FmmHandle *> shared;
  for (int i = 0; i < nh; ++i) {
    FmmHandle *h = hs[i]->h;
    if (!tuned_batchable(h)) continue;
    if (!shared.empty() && (h->prm.use_vs != shared[0]->prm.use_vs ||
                            h->prm.kernel != shared[0]->prm.kernel ||
                            (h->prm.use_vs && h->prm.use_divfilt) !=
                                (shared[0]->prm.use_vs && shared[0]->prm.use_divfilt)))
      continue;
    shared.push_back(h);
  }
  struct Events {
    hipEvent_t m2l = nullptr, p2p = nullptr;
    ~Events() {
      if (m2l) (void)hipEventDestroy(m2l);
      if (p2p) (void)hipEventDestroy(p2p);
    }
  } ev;
  auto clear = [&]() {
    for (FmmHandle *h : shared) h->batched = false;
  };
  if (!shared.empty()) {
    DUST_HIP(hipEventCreateWithFlags(&ev.m2l, hipEventDisableTiming));
    DUST_HIP(hipEventCreateWithFlags(&ev.p2p, hipEventDisableTiming));
    const int rc = tuned_batch_launch(shared.data(), (int)shared.size(), shared[0]->stream, ev.m2l,
                                      ev.p2p);
    if (rc) {
      (void)sync_all();
      return rc;
    }
    for (FmmHandle *h : shared) h->batched = true;
  }
  int rc = 0;
  for (int i = 0; i < nh && rc == 0; ++i) {
    FmmHandle *h = hs[i]->h;
    if (h->batched) {
      if (hipStreamWaitEvent(h->stream, ev.m2l, 0) != hipSuccess) {
        set_error("hipStreamWaitEvent failed");
        rc = DUST_E_HIP;
        break;
      }
    }
    rc = multipole_phase_b(h);
    if (rc == 0 && h->batched && hipStreamWaitEvent(h->stream, ev.p2p, 0) != hipSuccess) {
      set_error("hipStreamWaitEvent failed");
      rc = DUST_E_HIP;
    }
    if (rc == 0) rc = fmm_apply(h);
  }
  const int rs = sync_all();  // also before the events are destroyed
  clear();
  return rc ? rc : rs;
}

void *dust_fmm_stream(dust_fmm_t h) { return h ? (void *)h->h->stream : nullptr; }

int dust_fmm_graph_reserve(dust_fmm_t h, double slack) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return graph_reserve(h->h, slack);
}

int dust_fmm_pass_graph_dev(dust_fmm_t h, const int *n_dev, const double *cen, const double *dir,
                            const double *mag, const double *rv, double *vel, double *stretch) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return graph_enqueue(h->h, n_dev, cen, dir, mag, rv, vel, stretch);
}

int dust_fmm_graph_capture(dust_fmm_t h, const int *n_dev, const double *cen, const double *dir,
                           const double *mag, const double *rv, double *vel, double *stretch) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return graph_capture(h->h, n_dev, cen, dir, mag, rv, vel, stretch);
}

int dust_fmm_graph_launch(dust_fmm_t h) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return graph_launch(h->h);
}

int dust_fmm_graph_status(dust_fmm_t h, int *counts) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  return graph_status(h->h, counts);
}

int dust_fmm_synchronize(dust_fmm_t h) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  DUST_HIP(hipStreamSynchronize(h->h->stream));
  timer_collect(h->h);
  return 0;
}

int dust_fmm_set_tuning(dust_fmm_t h, const char *name, double value) {
  DUST_CHECK(h && name, DUST_E_ARG, "null argument");
  if (!strcmp(name, "seg_iters")) {  // both forms
    DUST_CHECK(value >= 64.0 && value <= 65536.0, DUST_E_ARG, "seg_iters must be in 64..65536");
    h->h->seg_iters = (int)value;
    return 0;
  }
  return tuned_set(h->h, name, value);
}

int dust_fmm_get_tuning(dust_fmm_t h, const char *name, double *value) {
  DUST_CHECK(h && name && value, DUST_E_ARG, "null argument");
  if (!strcmp(name, "seg_iters")) {
    *value = h->h->seg_iters;
    return 0;
  }
  return tuned_get(h->h, name, value);
}

// Times full passes over the particles already in the handle for a small grid
// of the performance knobs (P2P item length; M2L offsets per job in the tuned
// form) and keeps the fastest setting.  Knobs that change the numbers
// (near_factor) are left alone.  The handle's results are those of the last
// timed pass, not of the chosen setting: run a pass afterwards.
int dust_fmm_autotune(dust_fmm_t hh, int reps, double *best_ms) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  DUST_CHECK(h->n > 0, DUST_E_ARG, "autotune needs particles in the handle");
  reps = std::max(1, std::min(reps, 20));
  const int segs[] = {768, 1536, 3072};
  const int chunks[] = {7, 14, 21, 27};
  const bool tuned = h->tuned != nullptr;
  const int nchunk = tuned ? 4 : 1;
  hipEvent_t e0, e1;
  DUST_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("hipEventCreate failed");
    return DUST_E_HIP;
  }
  double best = INFINITY, chunk0 = 0.0;
  int best_seg = h->seg_iters, best_chunk = 0;
  if (tuned) DUST_TRY(tuned_get(h, "m2l_chunk", &chunk0));
  best_chunk = (int)chunk0;
  int rc = 0;
  for (int a = 0; a < 3 && !rc; ++a)
    for (int b = 0; b < nchunk && !rc; ++b) {
      h->seg_iters = segs[a];
      if (tuned) rc = tuned_set(h, "m2l_chunk", chunks[b]);
      double tmin = INFINITY;
      for (int r = 0; r <= reps && !rc; ++r) {  // r = 0 warms the buffers up
        int nl = 0;
        if ((rc = hipEventRecord(e0, h->stream) != hipSuccess ? DUST_E_HIP : 0)) break;
        if ((rc = fmm_sort(h, &nl)) || (rc = fmm_multipole(h)) || (rc = fmm_apply(h))) break;
        if (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          rc = DUST_E_HIP;
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) tmin = std::min(tmin, (double)ms);
      }
      if (!rc && tmin < best) {
        best = tmin;
        best_seg = segs[a];
        best_chunk = tuned ? chunks[b] : 0;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->seg_iters = best_seg;
  if (tuned && best_chunk > 0) DUST_TRY(tuned_set(h, "m2l_chunk", best_chunk));
  if (rc) return rc;
  if (best_ms) *best_ms = best;
  return 0;
}

int dust_fmm_set_timing(dust_fmm_t h, int enable) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  timer_collect(h->h);
  h->h->timer.on = enable != 0;
  for (int s = 0; s < DUST_FMM_NSTAGES; ++s) {
    h->h->timer.ms[s] = 0.0;
    h->h->timer.launches[s] = 0;
  }
  return 0;
}

int dust_fmm_stage_times(dust_fmm_t h, double *ms, int *launches) {
  DUST_CHECK(h, DUST_E_ARG, "null handle");
  timer_collect(h->h);
  for (int s = 0; s < DUST_FMM_NSTAGES; ++s) {
    if (ms) ms[s] = h->h->timer.ms[s];
    if (launches) launches[s] = h->h->timer.launches[s];
  }
  return 0;
}

int dust_fmm_leaves(dust_fmm_t hh, int *level, int *ijk, int *npart, int *members) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  DUST_CHECK(h->sorted, DUST_E_ARG, "sort_particles must be called first");
  TreeGeom G = make_geom(h);
  std::vector<int> lv(h->n_leaves), fs(h->nfine + 1), perm(h->ns);
  DUST_HIP(hipMemcpy(lv.data(), h->leaves.p, sizeof(int) * h->n_leaves, hipMemcpyDeviceToHost));
  DUST_HIP(hipMemcpy(fs.data(), h->fstart.p, sizeof(int) * (h->nfine + 1), hipMemcpyDeviceToHost));
  if (h->ns > 0)
    DUST_HIP(hipMemcpy(perm.data(), h->perm.p, sizeof(int) * h->ns, hipMemcpyDeviceToHost));
  int w = 0;
  for (int q = 0; q < h->n_leaves; ++q) {
    long g = lv[q];
    int l = 0;
    while (l + 1 < h->nlev && g >= h->lev[l + 1].off) ++l;
    long c = g - h->lev[l].off;
    int ix, iy, iz;
    cell_decode(G, l, c, ix, iy, iz);
    int sh = 3 * (h->nlev - 1 - l);
    int s = fs[c << sh], e = fs[(c + 1) << sh];
    if (level) level[q] = l + 1;
    if (ijk) { ijk[3 * q] = ix + 1; ijk[3 * q + 1] = iy + 1; ijk[3 * q + 2] = iz + 1; }
    if (npart) npart[q] = e - s;
    if (members)
      for (int t = s; t < e; ++t) members[w++] = perm[t];
  }
  return 0;
}

int dust_fmm_cell_index(dust_fmm_t hh, int *ijk) {
  DUST_CHECK(hh && ijk, DUST_E_ARG, "null argument");
  FmmHandle *h = hh->h;
  DUST_CHECK(h->sorted, DUST_E_ARG, "sort_particles must be called first");
  TreeGeom G = make_geom(h);
  std::vector<unsigned> ks(h->ns);
  std::vector<int> perm(h->ns);
  if (h->n == 0) return 0;
  std::fill(ijk, ijk + 3 * (size_t)h->n, 0);  // particles freed by a redistribution
  if (h->ns == 0) return 0;
  DUST_HIP(hipMemcpy(ks.data(), h->key_s.p, sizeof(unsigned) * h->ns, hipMemcpyDeviceToHost));
  DUST_HIP(hipMemcpy(perm.data(), h->perm.p, sizeof(int) * h->ns, hipMemcpyDeviceToHost));
  for (int t = 0; t < h->ns; ++t) {
    int ix, iy, iz;
    cell_decode(G, h->nlev - 1, (long)ks[t], ix, iy, iz);
    int q = perm[t];
    ijk[3 * q] = ix + 1; ijk[3 * q + 1] = iy + 1; ijk[3 * q + 2] = iz + 1;
  }
  return 0;
}

int dust_fmm_work(dust_fmm_t hh, double *p2p_pairs, double *m2l_pairs) {
  DUST_CHECK(hh, DUST_E_ARG, "null handle");
  FmmHandle *h = hh->h;
  DUST_CHECK(h->sorted, DUST_E_ARG, "sort_particles must be called first");
  TreeGeom G = make_geom(h);
  std::vector<int> cnt(h->n_leaves), lv(h->n_leaves), fs(h->nfine + 1);
  std::vector<int2> rng((size_t)h->n_leaves * h->p2p_cap);
  std::vector<unsigned char> fl(h->ncells_tot);
  if (h->n_leaves > 0) {
    DUST_HIP(hipMemcpy(cnt.data(), h->p2p_cnt.p, sizeof(int) * h->n_leaves, hipMemcpyDeviceToHost));
    DUST_HIP(hipMemcpy(lv.data(), h->leaves.p, sizeof(int) * h->n_leaves, hipMemcpyDeviceToHost));
    DUST_HIP(hipMemcpy(rng.data(), h->p2p_rng.p, sizeof(int2) * rng.size(), hipMemcpyDeviceToHost));
  }
  DUST_HIP(hipMemcpy(fs.data(), h->fstart.p, sizeof(int) * (h->nfine + 1), hipMemcpyDeviceToHost));
  DUST_HIP(hipMemcpy(fl.data(), h->flags.p, h->ncells_tot, hipMemcpyDeviceToHost));
  double pp = 0.0;
  for (int q = 0; q < h->n_leaves; ++q) {
    const int2 *r = &rng[(size_t)q * h->p2p_cap];
    double src = 0.0;
    for (int t = 0; t + 1 < cnt[q]; ++t) src += r[t].y - r[t].x;
    double np = r[cnt[q] - 1].y - r[cnt[q] - 1].x;
    pp += np * (src + np - 1.0);
  }
  double mp = 0.0;
  for (int l = 0; l < h->nlev; ++l) {
    const Level &L = h->lev[l];
    for (long c = 0; c < L.ncells; ++c) {
      if (!fl[L.off + c]) continue;
      int ix, iy, iz;
      cell_decode(G, l, c, ix, iy, iz);
      if (l == 0) {
        for (long s = 0; s < L.ncells; ++s) {
          int sx, sy, sz;
          cell_decode(G, 0, s, sx, sy, sz);
          if ((std::abs(sx - ix) > 1 || std::abs(sy - iy) > 1 || std::abs(sz - iz) > 1) && fl[L.off + s])
            mp += 1.0;
        }
        continue;
      }
      int px = ix >> 1, py = iy >> 1, pz = iz >> 1;
      for (int kc = -1; kc <= 1; ++kc)
        for (int jc = -1; jc <= 1; ++jc)
          for (int ic = -1; ic <= 1; ++ic) {
            if (!ic && !jc && !kc) continue;
            int a = px + ic, b = py + jc, d = pz + kc;
            const Level &P = h->lev[l - 1];
            if (a < 0 || b < 0 || d < 0 || a >= P.ncl[0] || b >= P.ncl[1] || d >= P.ncl[2]) continue;
            for (int p = 0; p < 8; ++p) {
              int sx = 2 * a + (p & 1), sy = 2 * b + ((p >> 1) & 1), sz = 2 * d + ((p >> 2) & 1);
              if (std::abs(sx - ix) <= 1 && std::abs(sy - iy) <= 1 && std::abs(sz - iz) <= 1) continue;
              if (fl[L.off + cell_encode(G, l, sx, sy, sz)]) mp += 1.0;
            }
          }
    }
  }
  if (p2p_pairs) *p2p_pairs = pp;
  if (m2l_pairs) *m2l_pairs = mp;
  return 0;
}

int dust_vortpart_direct(int n, const double *cen, const double *dir, const double *mag,
                         const double *rv, int use_vs, int use_vd, double nu_inf,
                         const double u_inf[3], double *vel, double *stretch) {
  return dust_vortpart_direct_k(n, cen, dir, mag, rv, use_vs, use_vd, nu_inf, u_inf,
                                DUST_KERNEL_ROSENHEAD, vel, stretch);
}

int dust_vortpart_direct_k(int n, const double *cen, const double *dir, const double *mag,
                           const double *rv, int use_vs, int use_vd, double nu_inf,
                           const double u_inf[3], int kernel, double *vel, double *stretch) {
  DUST_CHECK(n >= 0, DUST_E_ARG, "negative particle count");
  DUST_CHECK(kernel == DUST_KERNEL_ROSENHEAD || kernel == DUST_KERNEL_GAUSSIAN, DUST_E_ARG,
             "kernel must be DUST_KERNEL_ROSENHEAD or DUST_KERNEL_GAUSSIAN");
  if (n == 0) return 0;
  DUST_CHECK(cen && dir && mag && rv && vel, DUST_E_ARG, "null argument");
  double *d = nullptr;
  size_t tot = (size_t)n * 14;
  DUST_HIP(hipMalloc(&d, sizeof(double) * tot));
  double *dc = d, *dd = d + 3 * n, *dm = d + 6 * n, *dr = d + 7 * n, *dv = d + 8 * n,
         *ds = d + 11 * n;
  int rc = 0;
  if (hipMemcpy(dc, cen, sizeof(double) * 3 * n, hipMemcpyHostToDevice) ||
      hipMemcpy(dd, dir, sizeof(double) * 3 * n, hipMemcpyHostToDevice) ||
      hipMemcpy(dm, mag, sizeof(double) * n, hipMemcpyHostToDevice) ||
      hipMemcpy(dr, rv, sizeof(double) * n, hipMemcpyHostToDevice)) {
    set_error("upload failed");
    rc = DUST_E_HIP;
  }
  if (!rc) {
    double u0 = u_inf ? u_inf[0] : 0.0, u1 = u_inf ? u_inf[1] : 0.0, u2 = u_inf ? u_inf[2] : 0.0;
    k_direct<<<ceil_div(n, kEvalThreads), kEvalThreads>>>(n, dc, dd, dm, dr, use_vs, use_vd,
                                                          nu_inf, u0, u1, u2, kernel, dv, ds);
    if (hipMemcpy(vel, dv, sizeof(double) * 3 * n, hipMemcpyDeviceToHost) ||
        (stretch && hipMemcpy(stretch, ds, sizeof(double) * 3 * n, hipMemcpyDeviceToHost))) {
      set_error("direct evaluation failed: %s", hipGetErrorString(hipGetLastError()));
      rc = DUST_E_HIP;
    }
  }
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
