// Tuned evaluation form of the FMM on gfx950.
//
//  * M2L on the matrix cores: for each level and each octant of the target
//    inside its parent, every active target has the same 189 interaction
//    offsets (children of the parent's neighbours that are not neighbours,
//    mod_octree.f90:441-476).  The translation of M2L_multipole
//    (mod_multipole.f90:282-356) for one offset is a fixed matrix
//        [b(:,m); c(:,:,m)] += (-1)^|m| sum_n mult(m,n) cross(D(m+n) | Dc(:,i,m+n), a(:,n))
//    of 12 n_mon rows by 3 n_mon columns, so a whole (level, octant) group is
//    one GEMM  Y(12 n_mon x targets) += [W_o1 ... W_o189] x [A_o1; ...; A_o189]
//    with K = 189 x 3 n_mon.  The operators W are precomputed per (level,
//    offset) in FP32 and streamed in MFMA fragment order, the B operand
//    gathers the FP32 source multipoles, and v_mfma_f32_16x16x4f32
//    accumulates in FP32; the result is added into the FP64 local
//    expansions.
//  * P2P in FP32: coordinates relative to the target leaf centre, rsqrt on
//    the transcendental unit, per-tile FP32 partial sums folded into FP64
//    accumulators.  Pairs closer than the source core radius (times a
//    factor) are recomputed in FP64 for the whole wavefront.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fmm_eval.h"

namespace dust {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kOct = 189;  // interaction offsets per target octant

struct M2LJob {
  int lvl, oct, t0, j0, rb, nt;  // nt = targets in the group
  long grp_off;                  // first grouped target of (lvl, oct)
  long kd_base;
  int kdl0, kdl1, kdl2, pad;
};

struct TunedState {
  DevBuf<float> a32;        // FP32 copy of the multipole coefficients, (3 n_mon) per slot
  DevBuf<int> grp;          // target slots grouped by (level, octant)
  DevBuf<int> src;          // per grouped target: 189 source slots (-1 = none)
  DevBuf<long> pos;         // per active cell of a level: its grouped entry
  DevBuf<unsigned long long> gkey;  // group sort keys of the targets, in and sorted
  DevBuf<int> gidx;         // target numbers, in and sorted
  DevBuf<int> gstart;       // [kMaxLevels * 8] first sorted target of every group
  DevBuf<int> gcount;       // [kMaxLevels * 8] group sizes (device)
  DevBuf<int> offs;         // [8][189][3] offsets of each octant
  DevBuf<float> W;          // M2L operators of every level and offset, fragment order
  int m2l_chunk = 21;       // offsets per M2L job
  DevBuf<unsigned char> jobs;  // M2L job list of the current pass
  HostBuf<M2LJob> host_jobs;   // page-locked source of the job upload
  HostBuf<int> hcount;          // [kMaxLevels * 8] group sizes (page-locked)
  long gbase[kMaxLevels * 8];  // first entry of every group in grp
  double near_factor = 1.0;
  // graph form: group entries placed by level capacities, the job list built
  // on the device, the persistent M2L's arguments and its 8 queue counters
  long ggbase[kMaxLevels * 8];
  DevBuf<unsigned char> gjobs, gargs;
  DevBuf<int> gq, gtab;
  unsigned gm2l_grid = 0;
};

static TunedState *TS(FmmHandle *h) { return (TunedState *)h->tuned; }
static const TunedState *TS(const FmmHandle *h) { return (const TunedState *)h->tuned; }
static int build_w(FmmHandle *h);

int tuned_prepare(FmmHandle *h) {
  TunedState *t = new TunedState();
  h->tuned = t;
  // offsets of every octant, in (k, j, i) order of the parent's neighbours
  std::vector<int> offs(8 * kOct * 3);
  for (int g = 0; g < 8; ++g) {
    int gx = g & 1, gy = (g >> 1) & 1, gz = (g >> 2) & 1;
    int w = 0;
    for (int dk = -2 - gz; dk <= 3 - gz; ++dk)
      for (int dj = -2 - gy; dj <= 3 - gy; ++dj)
        for (int di = -2 - gx; di <= 3 - gx; ++di) {
          if (std::abs(di) <= 1 && std::abs(dj) <= 1 && std::abs(dk) <= 1) continue;
          offs[(g * kOct + w) * 3 + 0] = di;
          offs[(g * kOct + w) * 3 + 1] = dj;
          offs[(g * kOct + w) * 3 + 2] = dk;
          ++w;
        }
    if (w != kOct) {
      set_error("internal: octant offset count %d", w);
      return DUST_E_HIP;
    }
  }
  DUST_TRY(t->offs.reserve(offs.size()));
  DUST_HIP(hipMemcpy(t->offs.p, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice));
  DUST_TRY(t->gcount.reserve(kMaxLevels * 8));
  DUST_TRY(t->hcount.reserve(kMaxLevels * 8));
  return build_w(h);
}

// Knobs of the tuned form (set by the autotuner or the caller).
int tuned_get(FmmHandle *h, const char *name, double *value) {
  TunedState *t = TS(h);
  DUST_CHECK(t, DUST_E_ARG, "the handle was not created with DUST_FMM_TUNED");
  if (!strcmp(name, "near_factor")) {
    *value = t->near_factor;
    return 0;
  }
  if (!strcmp(name, "m2l_chunk")) {
    *value = t->m2l_chunk;
    return 0;
  }
  set_error("unknown tuning knob '%s'", name);
  return DUST_E_ARG;
}

int tuned_set(FmmHandle *h, const char *name, double value) {
  TunedState *t = TS(h);
  DUST_CHECK(t, DUST_E_ARG, "the handle was not created with DUST_FMM_TUNED");
  if (!strcmp(name, "near_factor")) {
    DUST_CHECK(value >= 0.0, DUST_E_ARG, "near_factor must be >= 0");
    t->near_factor = value;
    return 0;
  }
  if (!strcmp(name, "m2l_chunk")) {
    DUST_CHECK(value >= 1.0 && value <= kOct, DUST_E_ARG, "m2l_chunk must be in 1..189");
    t->m2l_chunk = (int)value;
    return 0;
  }
  set_error("unknown tuning knob '%s'", name);
  return DUST_E_ARG;
}

void tuned_release(FmmHandle *h) {
  TunedState *t = TS(h);
  if (!t) return;
  t->a32.release();
  t->grp.release();
  t->src.release();
  t->pos.release();
  t->gcount.release();
  t->offs.release();
  t->W.release();
  t->jobs.release();
  t->host_jobs.release();
  t->hcount.release();
  delete t;
  h->tuned = nullptr;
}

// Targets of an M2L group (level, octant) are ordered by the pattern of
// active cells around them, so that the 16 targets of an MFMA column tile
// tend to have sources at the same offsets (tiles with no source at an
// offset are skipped).  The interaction offsets of a target are the
// children of its parent's neighbours, and the children of a cell are all
// active when it is a branch and all inactive otherwise, so the sources of
// a target are fixed by which of the 27 cells around its parent are
// branches.  Key of active cell a1 + q (levels >= 1): level (bits 30..33),
// octant (27..29) and 27 bits, bit 9 (px + 1) + 3 (py + 1) + (pz + 1) set
// when the cell (px, py, pz) away from the parent is a branch.  Counted on
// the 1M-particle helix wake, issued / useful M2L work is 1.85 with the
// targets in Morton order and 1.43 in this order.
constexpr int kKeyBits = 34;
__global__ void k_group_key(TreeGeom G, const int *__restrict__ active, int a1, int n,
                            const unsigned char *__restrict__ flags, unsigned long long *key,
                            int *idx, int *gcount) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (G.dc) {
    a1 = G.dc[kGcActStart + 1];
    n = g_stop(G) ? 0 : G.dc[kGcActive] - a1;
  }
  if (q >= n) return;
  const long g = active[a1 + q];
  const int lvl = lvl_of(G, g);
  const long c = g - G.off[lvl];
  int ix, iy, iz;
  cell_decode(G, lvl, c, ix, iy, iz);
  const int oc = (ix & 1) + 2 * (iy & 1) + 4 * (iz & 1);
  const int pl = lvl - 1, qx = ix >> 1, qy = iy >> 1, qz = iz >> 1;
  unsigned pat = 0;
  for (int px = -1; px <= 1; ++px)
    for (int py = -1; py <= 1; ++py)
      for (int pz = -1; pz <= 1; ++pz) {
        const int sx = qx + px, sy = qy + py, sz = qz + pz;
        if (sx < 0 || sy < 0 || sz < 0 || sx >= G.ncl[pl][0] || sy >= G.ncl[pl][1] ||
            sz >= G.ncl[pl][2])
          continue;
        if (flags[G.off[pl] + cell_encode(G, pl, sx, sy, sz)] == kBranch)
          pat |= 1u << (9 * (px + 1) + 3 * (py + 1) + (pz + 1));
      }
  key[q] = ((unsigned long long)lvl << 30) | ((unsigned long long)oc << 27) | pat;
  idx[q] = q;
  atomicAdd(&gcount[lvl * 8 + oc], 1);
}

// First sorted target of every group: exclusive prefix of the group sizes
// in (level, octant) order, the order of the sorted keys.
__global__ void k_group_starts(const int *__restrict__ gcount, int *gstart) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int s = 0;
  for (int e = 0; e < kMaxLevels * 8; ++e) {
    gstart[e] = s;
    s += gcount[e];
  }
}

struct GroupBase {
  long b[kMaxLevels * 8];  // first entry of every group in grp
};

// Sorted target i: entry gbase + (i - first sorted target of its group).
__global__ void k_group_place(int a1, int n, const int *__restrict__ active,
                              const unsigned long long *__restrict__ key,
                              const int *__restrict__ idx, const int *__restrict__ gstart,
                              GroupBase gb, const int *__restrict__ slot, int *grp, long *pos,
                              const int *__restrict__ dc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (dc) {
    a1 = dc[kGcActStart + 1];
    n = dc[kGcOvf] ? 0 : dc[kGcActive] - a1;
  }
  if (i >= n) return;
  const int e8 = (int)(key[i] >> 27);  // level * 8 + octant
  const int q = idx[i];
  const long e = gb.b[e8] + (i - gstart[e8]);
  grp[e] = slot[active[a1 + q]];
  pos[a1 + q] = e;
}

// The 189 source slots of every grouped target (-1 = no active cell there):
// one thread per (target, offset), neighbouring threads on neighbouring
// offsets of one target.
__global__ void k_group_src(TreeGeom G, int lvl, const int *__restrict__ active, int a0, int na,
                            const int *__restrict__ slot, const unsigned char *__restrict__ flags,
                            const int *__restrict__ offs, const long *__restrict__ pos, int *src) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (G.dc) {
    a0 = G.dc[kGcActStart + lvl];
    na = g_stop(G) ? 0 : G.dc[kGcActStart + lvl + 1] - a0;
    pos += a0;
  }
  if (t >= (long)na * kOct) return;
  const int q = (int)(t / kOct), j = (int)(t % kOct);
  long c = active[a0 + q] - G.off[lvl];
  int ix, iy, iz;
  cell_decode(G, lvl, c, ix, iy, iz);
  const int oc = (ix & 1) + 2 * (iy & 1) + 4 * (iz & 1);
  const int sx = ix + offs[(oc * kOct + j) * 3], sy = iy + offs[(oc * kOct + j) * 3 + 1],
            sz = iz + offs[(oc * kOct + j) * 3 + 2];
  int s = -1;
  if (sx >= 0 && sy >= 0 && sz >= 0 && sx < G.ncl[lvl][0] && sy < G.ncl[lvl][1] &&
      sz < G.ncl[lvl][2]) {
    long gs = G.off[lvl] + cell_encode(G, lvl, sx, sy, sz);
    if (flags[gs]) s = slot[gs];
  }
  src[pos[q] * kOct + j] = s;
}

__global__ void k_to_f32(long n, const double *__restrict__ a, float *b) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// Same over the multipoles of the active cells counted on the device.
__global__ void k_to_f32_dc(long per, const double *__restrict__ a, float *b,
                            const int *__restrict__ dc) {
  const long n = dc[kGcOvf] ? 0 : (long)dc[kGcActive] * per;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
// M2L as one GEMM per (level, octant)
//
// Operator W_o (rows: rpm per monomial m, rpm = 12 with stretching, 3 without;
// columns: 3 per monomial n, the components of a(:,n)) is stored once per
// handle for every level and offset in MFMA fragment order: for row tile rt,
// k step ks and lane l, element (row 16 rt + l % 16, col 4 ks + l / 16), so
// that a wavefront loads one fragment with one coalesced 256-byte access.
// ---------------------------------------------------------------------------
constexpr int kNT = 64;       // targets per block (4 column tiles of 16)
constexpr int kRTW = 7;       // row tiles per wavefront
constexpr int kRTB = 4 * kRTW;

struct WGeom {
  int rpm, nrows, K, rt, ks;  // rows per monomial, rows, K, row tiles, k steps
  int rtp;                    // row tiles stored: rt rounded up to whole row blocks
  long per_off;               // floats per offset
};

// The operator of one offset is stored as rtp x ks fragments; ks is even and
// rows past nrows / columns past K are zero, so every fragment load of a
// row block is in bounds and the k loop runs in pairs.
static WGeom wgeom(const FmmHandle *h) {
  WGeom w;
  w.rpm = h->prm.use_vs ? 12 : 3;
  w.nrows = w.rpm * h->nmon;
  w.K = 3 * h->nmon;
  w.rt = (w.nrows + 15) / 16;
  w.rtp = ceil_div(w.rt, kRTB) * kRTB;
  w.ks = 2 * ceil_div(w.K, 8);
  w.per_off = (long)w.rtp * w.ks * 64;
  return w;
}

// Levi-Civita symbol for distinct indices in 0..2.
__device__ inline double eps3(int a, int b) { return ((b - a + 3) % 3 == 1) ? 1.0 : -1.0; }

// W_o(row, col) = (-1)^|m| mult(m,n) eps(o3, c', e) T_(c'[,i])(m+n), with
// row = (m, comp), col = (n, e): comp < 3 gives b(comp, m) from D(c');
// comp = 3 + 3 j + i gives c(i, j, m) from Dc(c', i) (cross product index j).
__global__ void k_build_w(WGeom w, int nmon, int nU, long ntab, const double *__restrict__ kd,
                          const double *__restrict__ multT, const int *__restrict__ mnT,
                          const int *__restrict__ pwr, float *W) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ntab * w.per_off) return;
  long tab = q / w.per_off;
  long r = q % w.per_off;
  int lane = (int)(r & 63);
  long fr = r >> 6;
  int ks = (int)(fr % w.ks), rt = (int)(fr / w.ks);
  int row = rt * 16 + (lane & 15), col = ks * 4 + (lane >> 4);
  float v = 0.f;
  if (row < w.nrows && col < w.K) {
    int m = row / w.rpm, comp = row % w.rpm;
    int n = col / 3, e = col % 3;
    int o3 = comp < 3 ? comp : (comp - 3) / 3;
    if (o3 != e) {
      int cp = 3 - o3 - e;
      int id = mnT[n * nmon + m];
      double mu = multT[n * nmon + m];
      int pw = pwr[3 * m] + pwr[3 * m + 1] + pwr[3 * m + 2];
      int trow = comp < 3 ? cp : 3 + cp + 3 * ((comp - 3) % 3);
      double sg = (pw & 1) ? -1.0 : 1.0;
      v = (float)(sg * eps3(o3, cp) * mu * kd[tab * 12 * (long)nU + (long)trow * nU + id]);
    }
  }
  W[q] = v;
}

// M2L operators of all levels and offsets (level 1 uses the FP64 kernel; its
// tables are built too so that every table index is valid).
static int build_w(FmmHandle *h) {
  TunedState *t = TS(h);
  WGeom w = wgeom(h);
  long nf = h->kd_tables * w.per_off;
  DUST_TRY(t->W.reserve(nf));
  k_build_w<<<ceil_div(nf, 256), 256, 0, h->stream>>>(w, h->nmon, h->nU, h->kd_tables, h->d_kd.p,
                                                     h->d_multT.p, h->d_mnT.p, h->d_pwr.p, t->W.p);
  DUST_HIP(hipGetLastError());
  DUST_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

// One job = (level, octant, 64 targets, a chunk of offsets, a block of
// row tiles); all jobs of a pass run in one launch and add their partial
// results into the FP64 local expansions with atomics (M2L only reads the
// multipoles, so every level can be translated before the L2L chain).


struct M2LArgs {
  WGeom w;
  int nmon, jc;
  const M2LJob *jobs;
  const int *grp, *src, *offs;
  const float *W;
  const float *a32;
  double *mp_b, *mp_c;
};

// One k step of a job: B fragments of the 64 targets from LDS, then the MFMAs
// of the wave's row tiles for the column tiles that have sources.
__device__ __forceinline__ void m2l_step(f4 (&acc)[kRTW][4], const float (&af)[kRTW],
                                         const float *sB, int ld, int ks, int lane, int bmask) {
  float bf[4];
  const int kk = ks * 4 + (lane >> 4);
#pragma unroll
  for (int b = 0; b < 4; ++b) bf[b] = sB[kk * ld + b * 16 + (lane & 15)];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    if (!(bmask & (1 << b))) continue;
#pragma unroll
    for (int a = 0; a < kRTW; ++a)  // padded row tiles hold zeros; their rows are not stored
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
  }
}

// One job of the M2L GEMMs (J.t0 < J.nt) by the 256 threads of the
// workgroup; sB, sSrc, sMask are the workgroup's LDS.
__device__ __forceinline__ void m2l_job(const M2LArgs &P, const M2LJob &J, float *sB, int *sSrc,
                                        int &sMask) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // the wave's first row tile, and its row tiles with rows < nrows
  const int rt0 = J.rb * kRTB + __builtin_amdgcn_readfirstlane(wv) * kRTW;
  const int nrt = max(0, min(kRTW, P.w.rt - rt0));
  const int K = P.w.K, KS = P.w.ks, ld = kNT + 4;
  const int *grp = P.grp + J.grp_off;
  const int *src = P.src + J.grp_off * kOct;
  f4 acc[kRTW][4];
#pragma unroll
  for (int a = 0; a < kRTW; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
  const int j1 = min(kOct, J.j0 + P.jc);
  // the source slot of this thread's target at the next offset, loaded one
  // offset ahead so that its latency hides behind the current offset's work
  const int tme = J.t0 + tid;
  int cur = (tid < kNT && tme < J.nt) ? src[(long)tme * kOct + J.j0] : -1;
  for (int j = J.j0; j < j1; ++j) {
    const int di = P.offs[(J.oct * kOct + j) * 3], dj = P.offs[(J.oct * kOct + j) * 3 + 1],
              dk = P.offs[(J.oct * kOct + j) * 3 + 2];
    const long tab =
        J.kd_base + (di + J.kdl0) +
        (long)(2 * J.kdl0 + 1) * ((dj + J.kdl1) + (long)(2 * J.kdl1 + 1) * (dk + J.kdl2));
    const float *Wo = P.W + tab * P.w.per_off;
    __syncthreads();
    int has = 0;
    if (tid < kNT) {
      sSrc[tid] = cur;
      has = cur >= 0;
      cur = (j + 1 < j1 && tme < J.nt) ? src[(long)tme * kOct + j + 1] : -1;
    }
    // column tiles (16 targets) with at least one source at this offset
    if (tid < 64) {
      const unsigned long long bal = __ballot(has);
      if (tid == 0) {
        int m = 0;
        for (int b = 0; b < 4; ++b) m |= ((bal >> (16 * b)) & 0xffffull) ? (1 << b) : 0;
        sMask = m;
      }
    }
    __syncthreads();
    const int bmask = sMask;
    if (bmask == 0) continue;  // no target has a source at this offset
    // gather the source multipoles of the 64 targets: sB[k][t], sixteen
    // loads per thread in flight at a time (clamped addresses; the missing
    // sources and the padding rows k >= K become zeros at the LDS store)
    {
      const int KS4 = KS * 4, tot = kNT * KS4;
      for (int q0 = 0; q0 < tot; q0 += 256 * 16) {
        float v[16];
        int dst[16];
        unsigned okm = 0u;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int q = q0 + tid + 256 * u;
          const int qq = min(q, tot - 1);
          const int tt = qq / KS4, k = qq - tt * KS4;
          const int s = sSrc[tt];
          v[u] = P.a32[(long)max(s, 0) * K + min(k, K - 1)];
          dst[u] = q < tot ? k * ld + tt : -1;
          if (q < tot && s >= 0 && k < K) okm |= 1u << u;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (dst[u] >= 0) sB[dst[u]] = ((okm >> u) & 1u) ? v[u] : 0.f;
      }
    }
    __syncthreads();
    if (nrt > 0) {
      // fragments of row tile a, k step ks at Wl[(a KS + ks) 64]; two register
      // sets alternate so the loads of one k step overlap the MFMAs of the other
      const float *Wl = Wo + ((long)rt0 * KS) * 64 + lane;
      float f0[kRTW], f1[kRTW];
#pragma unroll
      for (int a = 0; a < kRTW; ++a) f0[a] = Wl[((long)a * KS) * 64];
      for (int ks = 0; ks < KS; ks += 2) {
#pragma unroll
        for (int a = 0; a < kRTW; ++a) f1[a] = Wl[((long)a * KS + ks + 1) * 64];
        m2l_step(acc, f0, sB, ld, ks, lane, bmask);
        if (ks + 2 < KS) {
#pragma unroll
          for (int a = 0; a < kRTW; ++a) f0[a] = Wl[((long)a * KS + ks + 2) * 64];
        }
        m2l_step(acc, f1, sB, ld, ks + 1, lane, bmask);
      }
    }
  }
  // FP32 accumulator layout: lane holds column lane % 16, rows 4 (lane / 16) + i
  const int nm = P.nmon, rpm = P.w.rpm;
#pragma unroll
  for (int a = 0; a < kRTW; ++a) {
    if (a >= nrt) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int t = J.t0 + b * 16 + (lane & 15);
      if (t >= J.nt) continue;
      const long sl = grp[t];
      for (int i = 0; i < 4; ++i) {
        const int row = (rt0 + a) * 16 + 4 * (lane >> 4) + i;
        if (row >= P.w.nrows) continue;
        const int m = row / rpm, comp = row % rpm;
        const double v = (double)acc[a][b][i];
        if (comp < 3)
          atomicAdd(&P.mp_b[sl * 3 * nm + 3 * m + comp], v);
        else  // comp - 3 = 3 j + i holds c(i, j, m), stored at i + 3 j
          atomicAdd(&P.mp_c[sl * 9 * nm + 9 * m + (comp - 3)], v);
      }
    }
  }
}

// Workgroup b runs job b of one pass.
__global__ void __launch_bounds__(256) k_m2l_mfma(M2LArgs P) {
  extern __shared__ float sB[];  // [ks * 4][kNT + 4]
  __shared__ int sSrc[kNT];
  __shared__ int sMask;
  const M2LJob J = P.jobs[blockIdx.x];
  if (J.t0 >= J.nt) return;  // padding job
  m2l_job(P, J, sB, sSrc, sMask);
}

// The jobs of several passes in one persistent launch.  The job list holds 8
// queues interleaved (job k of queue x at k * 8 + x, each queue the jobs of
// the operator chunks dealt to XCD x; the case of a job in J.pad, its
// arguments in cases[J.pad]); workgroups are dispatched to the XCDs
// round-robin, so workgroup b runs on XCD b % 8 and takes the next job of
// queue b % 8 (counter qcnt[b % 8]) until the queue is empty: dynamic load
// balance inside an XCD, every operator chunk read into one L2.  Two
// workgroups per SIMD pair, as the one-pass kernel reaches.
__global__ void __launch_bounds__(256, 2) k_m2l_mfma_batch(const M2LJob *__restrict__ jobs,
                                                           const M2LArgs *__restrict__ cases,
                                                           int qlen, int *qcnt,
                                                           const int *__restrict__ dqlen) {
  extern __shared__ float sB[];  // [ks * 4][kNT + 4]
  __shared__ int sSrc[kNT];
  __shared__ int sMask;
  __shared__ int sJob;
  const int x = blockIdx.x & 7;
  if (dqlen) qlen = *dqlen;
  for (;;) {
    if (threadIdx.x == 0) sJob = atomicAdd(&qcnt[x], 1);
    __syncthreads();
    const int k = sJob;
    __syncthreads();  // read by every thread before thread 0 takes the next
    if (k >= qlen) break;
    const M2LJob J = jobs[(long)k * 8 + x];
    if (J.t0 >= J.nt) continue;  // padding job
    const M2LArgs P = cases[J.pad];
    m2l_job(P, J, sB, sSrc, sMask);
  }
}

__global__ void k_batch_clear(int *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// FP32 P2P with FP64 near pairs
// ---------------------------------------------------------------------------
// Near pair in FP64 (rsq and reciprocal refined by Newton steps instead of
// IEEE division): r = target - source, alpha_s = dir_s mag_s, core radius rv.
template <int GAUSS>
__device__ inline void near64(double r0, double r1, double r2, double al0, double al1, double al2,
                              double as0, double as1, double as2, double rv, int use_vs,
                              double v[3], double s[3], double *ru) {
  double cf, df;
  if (GAUSS) {
    gauss_cd(r0 * r0 + r1 * r1 + r2 * r2, rv, cf, df);
  } else {
    const double R2 = r0 * r0 + r1 * r1 + r2 * r2 + rv * rv;
    double y = __builtin_amdgcn_rsq(R2);
    y = y * (1.5 - 0.5 * R2 * y * y);
    y = y * (1.5 - 0.5 * R2 * y * y);
    const double y2 = y * y;
    cf = -(y2 * y);  // -1/R^3
    df = -3.0 * cf * y2;  // 3/R^5
  }
  v[0] += cf * (r1 * as2 - r2 * as1);
  v[1] += cf * (r2 * as0 - r0 * as2);
  v[2] += cf * (r0 * as1 - r1 * as0);
  if (use_vs) {
    const double vp0 = al1 * as2 - al2 * as1, vp1 = al2 * as0 - al0 * as2,
                 vp2 = al0 * as1 - al1 * as0;
    const double dv = r0 * vp0 + r1 * vp1 + r2 * vp2;
    s[0] -= cf * vp0 + df * r0 * dv;
    s[1] -= cf * vp1 + df * r1 * dv;
    s[2] -= cf * vp2 + df * r2 * dv;
    if (ru) {  // compute_rotu: -2 c alpha_s + d r x (r x alpha_s)
      const double i0 = r1 * as2 - r2 * as1, i1 = r2 * as0 - r0 * as2, i2 = r0 * as1 - r1 * as0;
      ru[0] += -2.0 * cf * as0 + df * (r1 * i2 - r2 * i1);
      ru[1] += -2.0 * cf * as1 + df * (r2 * i0 - r0 * i2);
      ru[2] += -2.0 * cf * as2 + df * (r0 * i1 - r1 * i0);
    }
  }
}

// FP32 P2P of one item (tuned form).  Sources are staged per wavefront as
// FP32 records (offset from the target leaf centre, r^2; alpha = dir*mag,
// near2 r^2) and FP64 records (offset, alpha, r).  Pairs with d^2 <= near2 r_s^2
// (inside the source core; the self pair is dropped by index) are left out
// of the FP32 sums and evaluated in
// FP64 after the tile, all lanes together (a per-lane mask of tile
// positions), so the rare FP64 work does not diverge the FP32 loop.  The
// FP32 sums of every tile are folded into FP64 accumulators.  A partial tile
// is padded with zero-strength sources far away, so the loop trip count is
// uniform.
//
// VS = 0: velocity; 1: velocity and stretching; 2: also the divergence
// filter's rotu, -2 c alpha_s + d r x (r x alpha_s) = alpha_s (3 t d^2 - 2 c)
// - 3 t (r.alpha_s) r with t = -d / 3 (the tf of the stretching sums).
// BATCH = 0: wavefront w of the grid evaluates item w of one pass (A0).
// BATCH = 1: the items of several passes in one persistent launch: every
// wavefront takes the next global item number from the counter *cnt until
// all are taken; item g belongs to the case c with off[c] <= g < off[c + 1]
// (off[ncase] = total) and is item g - off[c] of that case's pass (cases[c]).
// The body is written once, inside the kernel: as a called function it was
// compiled to a different register allocation and ran 15 % slower.
template <int VS, int GAUSS, int BATCH>
__global__ void __launch_bounds__(kWave * kEvalWaves)
    k_p2p_f32(EvalArgs A0, const EvalArgs *__restrict__ cases, const int *__restrict__ off,
              int ncase, int *cnt) {
  __shared__ f4 ssm[kEvalWaves][2 * kWave];
  __shared__ double sdm[kEvalWaves][8 * kWave];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  f4 *ss = ssm[wv];
  double *sd = sdm[wv];
  const int total = BATCH ? off[ncase]
                          : (A0.G.dc ? (g_stop(A0.G) ? 0 : A0.G.dc[kGcItems]) : A0.n_items);
  // the next item of the batch, taken by lane 0 for the wavefront
  auto next_item = [&]() {
    int v = 0;
    if (lane == 0) v = atomicAdd(cnt, 1);
    return __builtin_amdgcn_readfirstlane(v);
  };
  int gi = BATCH ? next_item() : blockIdx.x * kEvalWaves + wv;
  if (gi >= total) return;
  do {
    int cs = 0;
    if (BATCH) {
      int lo = 0, hi = ncase - 1;  // the last case with off[c] <= gi
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= gi) lo = mid;
        else hi = mid - 1;
      }
      cs = __builtin_amdgcn_readfirstlane(lo);
    }
    const EvalArgs &A = BATCH ? cases[cs] : A0;
    const int it = BATCH ? gi - off[cs] : gi;
    const EvalItem I = A.items[it];
    const EvalUnit U = A.units[I.unit];
    const long g = A.leaves[U.leaf];
    const int lvl = lvl_of(A.G, g);
    const long c = g - A.G.off[lvl];
    int ix, iy, iz;
    cell_decode(A.G, lvl, c, ix, iy, iz);
    const double cx = cell_center(A.G, lvl, 0, ix), cy = cell_center(A.G, lvl, 1, iy),
                 cz = cell_center(A.G, lvl, 2, iz);
    const int lpt = U.lpt, tl = lane / lpt, sub = lane - tl * lpt;
    const bool valid = tl < U.nt;
    const int t = U.t0 + (valid ? tl : 0);
    const double xr = A.px[t] - cx, yr = A.py[t] - cy, zr = A.pz[t] - cz;
    const double mg = A.pm[t];
    const double al0 = mg * A.dx[t], al1 = mg * A.dy[t], al2 = mg * A.dz[t];
    const float fx = (float)xr, fy = (float)yr, fz = (float)zr;
    const float fa0 = (float)al0, fa1 = (float)al1, fa2 = (float)al2;
    double pv[3] = {0, 0, 0}, ps[3] = {0, 0, 0}, pr[3] = {0, 0, 0};
    const double near2d = A.near2;  // staged per source as near2 r_s^2
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
          const double smg = A.pm[q], srv = A.pr[q];
          const double ox = A.px[q] - cx, oy = A.py[q] - cy, oz = A.pz[q] - cz;
          const double a0 = A.dx[q] * smg, a1 = A.dy[q] * smg, a2 = A.dz[q] * smg;
          // Rosenhead: r^2 of the core; Gaussian: 1 / (sqrt2 r)
          ss[2 * lane] = (f4){(float)ox, (float)oy, (float)oz,
                              GAUSS ? (float)(0.70710678118654752 / srv) : (float)(srv * srv)};
          ss[2 * lane + 1] = (f4){(float)a0, (float)a1, (float)a2, (float)(near2d * srv * srv)};
          double *d = sd + 8 * lane;
          d[0] = ox; d[1] = oy; d[2] = oz; d[3] = srv;
          d[4] = a0; d[5] = a1; d[6] = a2;
        } else {
          ss[2 * lane] = (f4){1.0e6f, 1.0e6f, 1.0e6f, 1.f};
          ss[2 * lane + 1] = (f4){0.f, 0.f, 0.f, 0.f};
        }
        wave_sync();
        float v0 = 0.f, v1 = 0.f, v2 = 0.f, w0 = 0.f, w1 = 0.f, w2 = 0.f;
        float S0 = 0.f, S1 = 0.f, S2 = 0.f;
        float T0 = 0.f, T1 = 0.f, T2 = 0.f, Q0 = 0.f, Q1 = 0.f, Q2 = 0.f;
        unsigned long long nmask = 0ull;
        // one source pair of the tile: s = sub + k lpt
        auto pair = [&](int s) {
          const f4 sp = ss[2 * s], sa = ss[2 * s + 1];
          const float r0 = fx - sp[0], r1 = fy - sp[1], r2 = fz - sp[2];
          const float a0 = sa[0], a1 = sa[1], a2 = sa[2];
          const float d2 = r0 * r0 + r1 * r1 + r2 * r2;
          // pairs inside the source core go to FP64, all but the self pair
          // (skipped by index, as the reference does: a distinct particle at
          // the same position still adds its stretching term)
          const bool skip = d2 <= sa[3];
          if (__builtin_expect(__any(skip), 0)) {
            if (skip && s0 + s != t) nmask |= 1ull << s;
          }
          // cf = c of kernel_coeffs; tf = -d / 3 (so that the stretching sum
          // below is shared by both cores: Rosenhead d = -3 cf / R^2)
          float cf, tf;
          if (GAUSS) {
            const float r = sqrtf(d2), x = r * sp[3];
            const float E = erff(x), ex = expf(-x * x);
            const float ir = 1.f / r, ir3 = ir * ir * ir;
            cf = skip ? 0.f : -(E - 1.1283791670955126f * x * ex) * ir3;
            tf = skip ? 0.f
                      : -(E - 0.18806319451591877f * ex * x * (6.f + 4.f * x * x)) * ir3 * ir * ir;
          } else {
            const float R2 = d2 + sp[3];
            const float inv = __builtin_amdgcn_rsqf(R2);
            const float inv2 = inv * inv;
            cf = skip ? 0.f : -(inv2 * inv);
            tf = cf * inv2;
          }
          const float c0 = r1 * a2 - r2 * a1, c1 = r2 * a0 - r0 * a2, c2 = r0 * a1 - r1 * a0;
          v0 += cf * c0;
          v1 += cf * c1;
          v2 += cf * c2;
          if (VS) {
            // stretching -[cf (al x a) + df r (r.(al x a))]: sum cf a is crossed
            // with al after the tile; r.(al x a) = -al.(r x a), df = -3 cf / R^2
            // (the factor 3 is applied after the tile)
            S0 += cf * a0;
            S1 += cf * a1;
            S2 += cf * a2;
            const float tt = tf * (fa0 * c0 + fa1 * c1 + fa2 * c2);
            w0 += tt * r0;
            w1 += tt * r1;
            w2 += tt * r2;
            if (VS == 2) {
              const float td = tf * d2, tq = tf * (r0 * a0 + r1 * a1 + r2 * a2);
              T0 += td * a0;
              T1 += td * a1;
              T2 += td * a2;
              Q0 += tq * r0;
              Q1 += tq * r1;
              Q2 += tq * r2;
            }
          }
        };
        // wave-uniform trip count: scalar loop control, two pairs per iteration
        const int nk = __builtin_amdgcn_readfirstlane((ns + lpt - 1) / lpt);
        int k = 0;
        for (; k + 2 <= nk; k += 2) {
          pair(sub + k * lpt);
          pair(sub + (k + 1) * lpt);
        }
        if (k < nk) pair(sub + k * lpt);
        if (VS) {
          w0 = -3.f * w0 - (fa1 * S2 - fa2 * S1);
          w1 = -3.f * w1 - (fa2 * S0 - fa0 * S2);
          w2 = -3.f * w2 - (fa0 * S1 - fa1 * S0);
        }
        if (VS == 2) {
          pr[0] += -2.f * S0 + 3.f * (T0 - Q0);
          pr[1] += -2.f * S1 + 3.f * (T1 - Q1);
          pr[2] += -2.f * S2 + 3.f * (T2 - Q2);
        }
        pv[0] += v0; pv[1] += v1; pv[2] += v2;
        ps[0] += w0; ps[1] += w1; ps[2] += w2;
        // the FP64 pairs of this tile, one per lane per step
        while (__any(nmask != 0ull)) {
          if (nmask) {
            const int s = __builtin_ctzll(nmask);
            nmask &= nmask - 1;
            const double *d = sd + 8 * s;
            near64<GAUSS>(xr - d[0], yr - d[1], zr - d[2], al0, al1, al2, d[4], d[5], d[6], d[3],
                          VS, pv, ps, VS == 2 ? pr : nullptr);
          }
        }
        wave_sync();
      }
      left -= e - (r.x + off);
    }
    for (int q = 0; q < 3; ++q) {
      pv[q] = lane_group_sum(pv[q], lpt);
      if (VS) ps[q] = lane_group_sum(ps[q], lpt);
      if (VS == 2) pr[q] = lane_group_sum(pr[q], lpt);
    }
    if (valid && sub == 0) {
      const double fi = 1.0 / (4.0 * kPi);
      double *o = &A.part[((long)it * kWave + tl) * (VS == 2 ? 9 : 6)];
      for (int q = 0; q < 3; ++q) {
        o[q] = pv[q] * fi;
        o[3 + q] = ps[q] * fi;
        if (VS == 2) o[6 + q] = pr[q] * fi;
      }
    }
    if (BATCH) gi = next_item();
  } while (BATCH && gi < total);
}

static size_t m2l_lds(const WGeom &w) { return sizeof(float) * (size_t)w.ks * 4 * (kNT + 4); }

// Per pass, after M2M: targets of every level grouped by octant with their
// 189 source slots, and the FP32 copy of the multipoles.
int tuned_pass_prepare(FmmHandle *h) {
  TunedState *t = TS(h);
  TreeGeom G = make_geom(h);
  DUST_HIP(hipMemsetAsync(t->gcount.p, 0, sizeof(int) * kMaxLevels * 8, h->stream));
  long tot = 0;
  for (int l = 0; l < h->nlev; ++l) {
    int na = h->act_start[l + 1] - h->act_start[l];
    for (int g = 0; g < 8; ++g) t->gbase[l * 8 + g] = tot + (long)g * na;
    tot += 8L * na;
  }
  DUST_TRY(t->grp.reserve(std::max(tot, 1L)));
  DUST_TRY(t->src.reserve(std::max(tot, 1L) * kOct));
  DUST_TRY(t->pos.reserve(std::max(h->n_active, 1)));
  // targets of every group in the order of their source patterns
  const int a1 = h->nlev > 1 ? h->act_start[1] : h->n_active;
  const int nt = h->n_active - a1;
  if (nt > 0) {
    DUST_TRY(t->gkey.reserve(2 * (size_t)nt));
    DUST_TRY(t->gidx.reserve(2 * (size_t)nt));
    DUST_TRY(t->gstart.reserve(kMaxLevels * 8));
    k_group_key<<<ceil_div(nt, 256), 256, 0, h->stream>>>(G, h->active.p, a1, nt, h->flags.p,
                                                          t->gkey.p, t->gidx.p, t->gcount.p);
    DUST_TRY(prim_sort_pairs_u64(h, t->gkey.p, t->gkey.p + nt, t->gidx.p, t->gidx.p + nt, nt,
                                 kKeyBits));
    k_group_starts<<<1, 64, 0, h->stream>>>(t->gcount.p, t->gstart.p);
    GroupBase gb;
    for (int e = 0; e < kMaxLevels * 8; ++e) gb.b[e] = e < h->nlev * 8 ? t->gbase[e] : 0;
    k_group_place<<<ceil_div(nt, 256), 256, 0, h->stream>>>(a1, nt, h->active.p, t->gkey.p + nt,
                                                            t->gidx.p + nt, t->gstart.p, gb,
                                                            h->slot.p, t->grp.p, t->pos.p, nullptr);
  }
  for (int l = 1; l < h->nlev; ++l) {
    int na = h->act_start[l + 1] - h->act_start[l];
    if (na == 0) continue;
    k_group_src<<<ceil_div((long)na * kOct, 256), 256, 0, h->stream>>>(
        G, l, h->active.p, h->act_start[l], na, h->slot.p, h->flags.p, t->offs.p,
        t->pos.p + h->act_start[l], t->src.p);
  }
  DUST_HIP(hipMemcpyAsync(t->hcount.p, t->gcount.p, sizeof(int) * kMaxLevels * 8,
                          hipMemcpyDeviceToHost, h->stream));
  // FP32 copy of the multipoles (P2M and M2M are complete at this point)
  long na3 = (long)h->n_active * 3 * h->nmon;
  DUST_TRY(t->a32.reserve(std::max(na3, 1L)));
  k_to_f32<<<ceil_div(na3, 256), 256, 0, h->stream>>>(na3, h->mp_a.p, t->a32.p);
  return 0;  // hcount is read after the caller's next synchronisation
}

// Chunks of M2L jobs: the targets of one (level, octant) group of case c
// against the offsets [j0, j0 + m2l_chunk), njobs = target tiles x row blocks.
struct M2LChunk {
  int c, lvl, g, j0, njobs;
};

static void m2l_chunks(const FmmHandle *h, int c, std::vector<M2LChunk> &out) {
  const TunedState *t = TS(h);
  const int nrb = ceil_div(wgeom(h).rt, kRTB);
  for (int lvl = 1; lvl < h->nlev; ++lvl)
    for (int g = 0; g < 8; ++g) {
      int nt = t->hcount.p[lvl * 8 + g];
      if (nt == 0) continue;
      for (int j0 = 0; j0 < kOct; j0 += t->m2l_chunk)
        out.push_back({c, lvl, g, j0, ceil_div(nt, kNT) * nrb});
    }
}

// The job list of the chunks (case c of a chunk is hs[c], recorded in
// M2LJob.pad) in `jobs`, njob entries.  Jobs that read the same operator chunk
// (case, level, octant, offset chunk) are given to one XCD: workgroups are
// dealt to the 8 XCDs round-robin by index, so the job of workgroup i comes
// from queue i % 8, and every chunk is fetched into one L2 instead of eight.
// Chunks go to the least loaded queue, largest first; short queues are padded
// with empty jobs.
static int m2l_job_list(FmmHandle *const *hs, std::vector<M2LChunk> &chunks,
                        HostBuf<M2LJob> &jobs, size_t *njob) {
  constexpr int kXcd = 8;
  std::stable_sort(chunks.begin(), chunks.end(),
                   [](const M2LChunk &a, const M2LChunk &b) { return a.njobs > b.njobs; });
  std::vector<std::vector<M2LJob>> queue(kXcd);
  std::vector<long> load(kXcd, 0);
  for (const M2LChunk &c : chunks) {
    const FmmHandle *h = hs[c.c];
    const TunedState *t = TS(h);
    const int nrb = ceil_div(wgeom(h).rt, kRTB);
    int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += c.njobs;
    const int nt = t->hcount.p[c.lvl * 8 + c.g];
    for (int t0 = 0; t0 < nt; t0 += kNT)
      for (int rb = 0; rb < nrb; ++rb) {
        M2LJob J;
        J.lvl = c.lvl; J.oct = c.g; J.t0 = t0; J.j0 = c.j0; J.rb = rb; J.nt = nt;
        J.grp_off = t->gbase[c.lvl * 8 + c.g];
        J.kd_base = h->lev[c.lvl].kd_base;
        J.kdl0 = h->lev[c.lvl].kdl[0]; J.kdl1 = h->lev[c.lvl].kdl[1]; J.kdl2 = h->lev[c.lvl].kdl[2];
        J.pad = c.c;
        queue[x].push_back(J);
      }
  }
  size_t qmax = 0;
  for (auto &q : queue) qmax = std::max(qmax, q.size());
  *njob = qmax * kXcd;
  if (qmax == 0) return 0;
  M2LJob empty{};
  empty.nt = 0;  // t0 >= nt: the job does nothing
  DUST_TRY(jobs.reserve(*njob));
  std::fill(jobs.p, jobs.p + *njob, empty);
  for (int x = 0; x < kXcd; ++x)
    for (size_t k = 0; k < queue[x].size(); ++k) jobs.p[k * kXcd + x] = queue[x][k];
  return 0;
}

static M2LArgs m2l_args(FmmHandle *h) {
  TunedState *t = TS(h);
  M2LArgs P;
  P.w = wgeom(h);
  P.nmon = h->nmon;
  P.jc = t->m2l_chunk;
  P.jobs = (const M2LJob *)t->jobs.p;
  P.grp = t->grp.p;
  P.src = t->src.p;
  P.offs = t->offs.p;
  P.W = t->W.p;
  P.a32 = t->a32.p;
  P.mp_b = h->mp_b.p;
  P.mp_c = h->mp_c.p;
  return P;
}

static int m2l_lds_attr(const void *kernel, size_t lds) {
  if (lds > 64 * 1024)
    DUST_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return 0;
}

// M2L of every level >= 2 (1-based) in one launch.
int tuned_m2l_all(FmmHandle *h) {
  TunedState *t = TS(h);
  const size_t lds = m2l_lds(wgeom(h));
  DUST_TRY(m2l_lds_attr((const void *)k_m2l_mfma, lds));
  std::vector<M2LChunk> chunks;
  m2l_chunks(h, 0, chunks);
  // the host list stays in the tuned state until the next pass (whose first
  // phase synchronises the stream), so the asynchronous upload needs no wait
  size_t njob = 0;
  FmmHandle *hs[1] = {h};
  DUST_TRY(m2l_job_list(hs, chunks, t->host_jobs, &njob));
  if (njob == 0) return 0;
  DUST_TRY(t->jobs.reserve(njob * sizeof(M2LJob)));
  DUST_HIP(hipMemcpyAsync(t->jobs.p, t->host_jobs.p, njob * sizeof(M2LJob),
                          hipMemcpyHostToDevice, h->stream));
  const M2LArgs P = m2l_args(h);
  k_m2l_mfma<<<(unsigned)njob, 256, lds, h->stream>>>(P);
  DUST_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Several tuned passes: their M2L jobs in one persistent launch and their P2P
// items in another (the shared work lists of many independent cases).
// ---------------------------------------------------------------------------
struct BatchBufs {
  DevBuf<int> cnt;  // [0..7] M2L queue counters, [8] P2P item counter
  HostBuf<M2LJob> hjobs;
  DevBuf<M2LJob> jobs;
  HostBuf<M2LArgs> hm2l;
  DevBuf<M2LArgs> m2l;
  HostBuf<EvalArgs> heval;
  DevBuf<EvalArgs> eval;
  HostBuf<int> hoff;
  DevBuf<int> off;
};
static BatchBufs g_batch;

void tuned_batch_release() {
  g_batch.cnt.release();
  g_batch.hjobs.release();
  g_batch.jobs.release();
  g_batch.hm2l.release();
  g_batch.m2l.release();
  g_batch.heval.release();
  g_batch.eval.release();
  g_batch.hoff.release();
  g_batch.off.release();
}

// A tuned pass joins the shared launches when its P2P runs in FP32 (no PSE).
bool tuned_batchable(const FmmHandle *h) {
  return h->tuned && !h->prm.use_vd && !h->prm.use_tv && h->n_active > 0;
}

// Workgroups of a persistent launch: as many as the CUs hold at once (by the
// kernel's occupancy), a multiple of the 8 XCDs.
static int persistent_grid(const void *kernel, int threads, size_t lds, unsigned *grid) {
  int dev = 0, cus = 0, per_cu = 0;
  DUST_HIP(hipGetDevice(&dev));
  DUST_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  DUST_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
  const long g = (long)std::max(cus, 1) * std::max(per_cu, 1);
  *grid = (unsigned)((g + 7) / 8 * 8);
  return 0;
}

// hs: passes whose sort and multipole_phase_a are complete (the caller has
// synchronised their streams), all tuned_batchable with the same use_vs and
// particle core.  Launches on st: the M2L of all of them (then records
// ev_m2l), then their P2P items (then ev_p2p).  The passes' own M2L and P2P
// are skipped (FmmHandle.batched).  (The two launches on two streams, side by
// side, measured no faster.)
int tuned_batch_launch(FmmHandle *const *hs, int nh, hipStream_t st, hipEvent_t ev_m2l,
                       hipEvent_t ev_p2p) {
  BatchBufs &B = g_batch;
  if (nh == 0) {
    DUST_HIP(hipEventRecord(ev_m2l, st));
    DUST_HIP(hipEventRecord(ev_p2p, st));
    return 0;
  }
  // ---- M2L ----
  std::vector<M2LChunk> chunks;
  size_t lds = 0;
  DUST_TRY(B.hm2l.reserve(nh));
  for (int c = 0; c < nh; ++c) {
    m2l_chunks(hs[c], c, chunks);
    B.hm2l.p[c] = m2l_args(hs[c]);
    lds = std::max(lds, m2l_lds(wgeom(hs[c])));
  }
  size_t njob = 0;
  DUST_TRY(m2l_job_list(hs, chunks, B.hjobs, &njob));
  DUST_TRY(B.cnt.reserve(9));
  k_batch_clear<<<1, 64, 0, st>>>(B.cnt.p, 9);
  if (njob > 0) {
    const void *kern = (const void *)k_m2l_mfma_batch;
    DUST_TRY(m2l_lds_attr(kern, lds));
    DUST_TRY(B.jobs.reserve(njob));
    DUST_TRY(B.m2l.reserve(nh));
    DUST_HIP(hipMemcpyAsync(B.jobs.p, B.hjobs.p, njob * sizeof(M2LJob), hipMemcpyHostToDevice,
                            st));
    DUST_HIP(hipMemcpyAsync(B.m2l.p, B.hm2l.p, nh * sizeof(M2LArgs), hipMemcpyHostToDevice, st));
    unsigned grid = 0;
    DUST_TRY(persistent_grid(kern, 256, lds, &grid));
    k_m2l_mfma_batch<<<grid, 256, lds, st>>>(B.jobs.p, B.m2l.p, (int)(njob / 8), B.cnt.p,
                                              nullptr);
    DUST_HIP(hipGetLastError());
  }
  DUST_HIP(hipEventRecord(ev_m2l, st));
  // ---- P2P ----
  DUST_TRY(B.heval.reserve(nh));
  DUST_TRY(B.hoff.reserve(nh + 1));
  long total = 0;
  for (int c = 0; c < nh; ++c) {
    EvalArgs &A = B.heval.p[c];
    make_eval_args(hs[c], &A);
    const double nf = TS(hs[c])->near_factor;
    A.near2 = nf * nf;
    B.hoff.p[c] = (int)total;
    total += A.n_items;
  }
  DUST_CHECK(total < 0x7fffffffL, DUST_E_CAPACITY, "%ld P2P items in one batch", total);
  B.hoff.p[nh] = (int)total;
  if (total > 0) {
    DUST_TRY(B.eval.reserve(nh));
    DUST_TRY(B.off.reserve(nh + 1));
    DUST_HIP(hipMemcpyAsync(B.eval.p, B.heval.p, nh * sizeof(EvalArgs), hipMemcpyHostToDevice,
                            st));
    DUST_HIP(hipMemcpyAsync(B.off.p, B.hoff.p, (nh + 1) * sizeof(int), hipMemcpyHostToDevice,
                            st));
    const bool vs = hs[0]->prm.use_vs, g = hs[0]->prm.kernel == DUST_KERNEL_GAUSSIAN;
    const bool dv = vs && hs[0]->prm.use_divfilt;
    using Kern = void (*)(EvalArgs, const EvalArgs *, const int *, int, int *);
    const Kern kern = dv ? (g ? k_p2p_f32<2, 1, 1> : k_p2p_f32<2, 0, 1>)
                      : vs ? (g ? k_p2p_f32<1, 1, 1> : k_p2p_f32<1, 0, 1>)
                           : (g ? k_p2p_f32<0, 1, 1> : k_p2p_f32<0, 0, 1>);
    unsigned grid = 0;
    DUST_TRY(persistent_grid((const void *)kern, kWave * kEvalWaves, 0, &grid));
    grid = std::min(grid, (unsigned)ceil_div(total, kEvalWaves));
    const EvalArgs &A0 = B.heval.p[0];  // not read by the batch instances
    kern<<<grid, kWave * kEvalWaves, 0, st>>>(A0, B.eval.p, B.off.p, nh, B.cnt.p + 8);
    DUST_HIP(hipGetLastError());
  }
  DUST_HIP(hipEventRecord(ev_p2p, st));
  return 0;
}

int tuned_eval(FmmHandle *h, EvalArgs &A) {
  TunedState *t = TS(h);
  if (h->prm.use_vd || h->prm.use_tv) return eval_fp64_launch(h, A);  // PSE runs in FP64
  A.near2 = t->near_factor * t->near_factor;
  if (A.n_items > 0) {
    const dim3 grid(ceil_div(A.n_items, kEvalWaves)), block(kWave * kEvalWaves);
    const bool g = A.kernel == DUST_KERNEL_GAUSSIAN;
    if (A.divf) {
      if (g) k_p2p_f32<2, 1, 0><<<grid, block, 0, h->stream>>>(A, nullptr, nullptr, 0, nullptr);
      else k_p2p_f32<2, 0, 0><<<grid, block, 0, h->stream>>>(A, nullptr, nullptr, 0, nullptr);
    } else if (A.use_vs) {
      if (g) k_p2p_f32<1, 1, 0><<<grid, block, 0, h->stream>>>(A, nullptr, nullptr, 0, nullptr);
      else k_p2p_f32<1, 0, 0><<<grid, block, 0, h->stream>>>(A, nullptr, nullptr, 0, nullptr);
    } else {
      if (g) k_p2p_f32<0, 1, 0><<<grid, block, 0, h->stream>>>(A, nullptr, nullptr, 0, nullptr);
      else k_p2p_f32<0, 0, 0><<<grid, block, 0, h->stream>>>(A, nullptr, nullptr, 0, nullptr);
    }
  }
  DUST_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Graph form of the tuned M2L: targets grouped as above with every count on
// the device, the job list built on the device in the order of
// m2l_job_list (chunks in (level, octant, offset chunk) order, stably sorted
// by job count, largest first; each to the least loaded of the 8 XCD queues;
// jobs of a chunk target tile by target tile, row blocks inside; short
// queues padded with empty jobs), then the persistent M2L over the queues.
// ---------------------------------------------------------------------------
constexpr int kGraphChunks = 2048;

// Dealing table of the graph form (int offsets in TunedState::gtab): chunk,
// queue, first queue position and jobs before, per dealt chunk; then the 8
// queue lengths, the number of dealt chunks, the job total and the longest
// queue.
constexpr int kTabOrd = 0, kTabQ = kGraphChunks, kTabSt = 2 * kGraphChunks,
              kTabPre = 3 * kGraphChunks, kTabQlen = 4 * kGraphChunks + 1,
              kTabNd = kTabQlen + 8, kTabNjob = kTabNd + 1, kTabQmax = kTabNjob + 1,
              kTabSize = kTabQmax + 1;

__global__ void __launch_bounds__(256)
    k_m2l_jobs(TreeGeom G, const int *__restrict__ gcount, int jc, int nrb, int job_cap, int *tab,
               int *dc) {
  __shared__ int cnj[kGraphChunks];  // jobs of every chunk
  __shared__ int sn[kGraphChunks];   // jobs of the r-th dealt chunk
  const int t = threadIdx.x;
  if (g_stop(G)) {
    if (t == 0) {
      dc[kGcQlen] = 0;
      tab[kTabQmax] = 0;
      tab[kTabNjob] = 0;
    }
    return;
  }
  const int nj = (kOct + jc - 1) / jc;
  const int nc = (G.nlev - 1) * 8 * nj;
  for (int q = t; q < nc; q += blockDim.x) {
    const int e = q / nj;
    const int nt = gcount[(1 + e / 8) * 8 + (e & 7)];
    cnj[q] = nt > 0 ? ((nt + kNT - 1) / kNT) * nrb : 0;
  }
  __syncthreads();
  // stable order by job count, largest first
  for (int q = t; q < nc; q += blockDim.x) {
    const int v = cnj[q];
    int r = 0;
    for (int p = 0; p < nc; ++p) {
      const int u = cnj[p];
      r += (u > v || (u == v && p < q)) ? 1 : 0;
    }
    tab[kTabOrd + r] = q;
    sn[r] = v;
  }
  __syncthreads();
  // each chunk to the least loaded queue (lowest index on ties)
  if (t == 0) {
    long load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int nd = nc, sum = 0;
    tab[kTabPre] = 0;
    for (int r = 0; r < nc; ++r) {
      const int v = sn[r];
      if (v == 0) {  // the rest hold no jobs (sorted)
        nd = r;
        break;
      }
      int x = 0;
#pragma unroll
      for (int y = 1; y < 8; ++y)
        if (load[y] < load[x]) x = y;
      tab[kTabQ + r] = x;
      tab[kTabSt + r] = (int)load[x];
      load[x] += v;
      sum += v;
      tab[kTabPre + r + 1] = sum;
    }
    long qm = 0;
    for (int y = 0; y < 8; ++y) {
      tab[kTabQlen + y] = (int)load[y];
      qm = load[y] > qm ? load[y] : qm;
    }
    if (qm * 8 > job_cap) {
      atomicOr(&dc[kGcOvf], 16);
      qm = 0;
      sum = 0;
    }
    tab[kTabNd] = nd;
    tab[kTabNjob] = sum;
    tab[kTabQmax] = (int)qm;
    dc[kGcQlen] = (int)qm;
  }
}

// The jobs of the dealt chunks (job j: the r-th dealt chunk with
// pre[r] <= j < pre[r + 1], searched in an LDS copy of pre) and the padding
// of the short queues, over the whole grid.
__global__ void __launch_bounds__(256)
    k_m2l_fill(TreeGeom G, const int *__restrict__ gcount, int jc, int nrb, GroupBase gb,
               const int *__restrict__ tab, M2LJob *jobs) {
  __shared__ int pre[kGraphChunks + 1];
  const int qmax = tab[kTabQmax];
  if (qmax == 0) return;
  const int nd = tab[kTabNd], njob = tab[kTabNjob];
  for (int r = threadIdx.x; r <= nd; r += blockDim.x) pre[r] = tab[kTabPre + r];
  __syncthreads();
  const int nj = (kOct + jc - 1) / jc;
  const int stride = gridDim.x * blockDim.x;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < njob; j += stride) {
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= j) lo = mid;
      else hi = mid - 1;
    }
    const int q = tab[kTabOrd + lo], k = j - pre[lo];
    const int e = q / nj, lvl = 1 + e / 8, g = e & 7;
    M2LJob J;
    J.lvl = lvl; J.oct = g; J.t0 = (k / nrb) * kNT; J.j0 = (q - e * nj) * jc; J.rb = k % nrb;
    J.nt = gcount[lvl * 8 + g];
    J.grp_off = gb.b[lvl * 8 + g];
    J.kd_base = G.kd_base[lvl];
    J.kdl0 = G.kdl[lvl][0]; J.kdl1 = G.kdl[lvl][1]; J.kdl2 = G.kdl[lvl][2];
    J.pad = 0;
    jobs[(long)(tab[kTabSt + lo] + k) * 8 + tab[kTabQ + lo]] = J;
  }
  for (long p = blockIdx.x * blockDim.x + threadIdx.x; p < 8L * qmax; p += stride) {
    const int y = (int)(p & 7), k = (int)(p >> 3);
    if (k < tab[kTabQlen + y]) continue;
    M2LJob J{};
    J.nt = 0;  // t0 >= nt: the job does nothing
    jobs[p] = J;
  }
}

// Buffers of the graph form, sized by the handle's graph capacities; the
// persistent M2L's arguments uploaded once.
int tuned_graph_reserve(FmmHandle *h) {
  TunedState *t = TS(h);
  GraphState &R = h->gr;
  const int nj = ceil_div(kOct, t->m2l_chunk);
  DUST_CHECK((h->nlev - 1) * 8 * nj <= kGraphChunks, DUST_E_ARG,
             "graph form: %d M2L chunks exceed %d (raise m2l_chunk)", (h->nlev - 1) * 8 * nj,
             kGraphChunks);
  long tot = 0;
  for (int l = 0; l < h->nlev; ++l) {
    for (int g = 0; g < 8; ++g) t->ggbase[l * 8 + g] = tot + (long)g * R.lev_cap[l];
    tot += 8L * R.lev_cap[l];
  }
  for (int e = h->nlev * 8; e < kMaxLevels * 8; ++e) t->ggbase[e] = 0;
  const long na = std::max(R.act_cap, 1);
  DUST_TRY(t->grp.reserve(std::max(tot, 1L)));
  DUST_TRY(t->src.reserve(std::max(tot, 1L) * kOct));
  DUST_TRY(t->pos.reserve(na));
  DUST_TRY(t->gkey.reserve(2 * na));
  DUST_TRY(t->gidx.reserve(2 * na));
  DUST_TRY(t->gstart.reserve(kMaxLevels * 8));
  DUST_TRY(t->a32.reserve(na * 3 * h->nmon));
  DUST_TRY(t->gq.reserve(8));
  DUST_TRY(t->gtab.reserve(kTabSize));
  // jobs: every group's target tiles times row blocks times chunks, with
  // one padding row per queue
  const int nrb = ceil_div(wgeom(h).rt, kRTB);
  long jobs = 0;
  for (int l = 1; l < h->nlev; ++l) jobs += 8L * (ceil_div(R.lev_cap[l], kNT) + 1) * nrb * nj;
  jobs = std::min(jobs, 8L * ((long)INT32_MAX / 16));
  R.job_cap = (int)std::max(jobs + 8L * nrb * nj * 8, 8L);
  DUST_TRY(t->gjobs.reserve((size_t)R.job_cap * sizeof(M2LJob)));
  DUST_TRY(t->gargs.reserve(sizeof(M2LArgs)));
  M2LArgs P = m2l_args(h);
  P.jobs = (const M2LJob *)t->gjobs.p;
  DUST_HIP(hipMemcpy(t->gargs.p, &P, sizeof(M2LArgs), hipMemcpyHostToDevice));
  DUST_TRY(m2l_lds_attr((const void *)k_m2l_mfma_batch, m2l_lds(wgeom(h))));
  DUST_TRY(persistent_grid((const void *)k_m2l_mfma_batch, 256, m2l_lds(wgeom(h)), &t->gm2l_grid));
  return 0;
}

void tuned_graph_buffers(const FmmHandle *h, std::vector<const void *> &v) {
  const TunedState *t = TS(h);
  const void *bufs[] = {t->grp.p, t->src.p, t->pos.p, t->gkey.p, t->gidx.p, t->gstart.p,
                        t->a32.p, t->gq.p, t->gjobs.p, t->gargs.p, t->gcount.p, t->offs.p, t->W.p,
                        t->gtab.p};
  for (const void *b : bufs) v.push_back(b);
}

// Tuned M2L of every level below the first, enqueued with device counts.
int tuned_graph_m2l(FmmHandle *h) {
  TunedState *t = TS(h);
  GraphState &R = h->gr;
  TreeGeom G = make_geom(h);
  int *dc = R.dc.p;
  const int cap = std::max(R.act_cap, 1);
  k_batch_clear<<<1, 128, 0, h->stream>>>(t->gcount.p, kMaxLevels * 8);
  k_batch_clear<<<1, 64, 0, h->stream>>>(t->gq.p, 8);
  k_group_key<<<ceil_div(cap, 256), 256, 0, h->stream>>>(G, h->active.p, 0, 0, h->flags.p,
                                                         t->gkey.p, t->gidx.p, t->gcount.p);
  DCount c{dc + kGcActive, dc + kGcActStart + 1, dc + kGcOvf, cap};
  DUST_TRY(gprim_sort_pairs_u64(h, t->gkey.p, t->gidx.p, t->gkey.p + cap, t->gidx.p + cap, c,
                                kKeyBits));
  k_group_starts<<<1, 64, 0, h->stream>>>(t->gcount.p, t->gstart.p);
  GroupBase gb;
  for (int e = 0; e < kMaxLevels * 8; ++e) gb.b[e] = t->ggbase[e];
  k_group_place<<<ceil_div(cap, 256), 256, 0, h->stream>>>(0, 0, h->active.p, t->gkey.p + cap,
                                                           t->gidx.p + cap, t->gstart.p, gb,
                                                           h->slot.p, t->grp.p, t->pos.p, dc);
  for (int l = 1; l < h->nlev; ++l) {
    if (R.lev_cap[l] == 0) continue;
    k_group_src<<<ceil_div((long)R.lev_cap[l] * kOct, 256L), 256, 0, h->stream>>>(
        G, l, h->active.p, 0, 0, h->slot.p, h->flags.p, t->offs.p, t->pos.p, t->src.p);
  }
  k_to_f32_dc<<<2048, 256, 0, h->stream>>>(3L * h->nmon, h->mp_a.p, t->a32.p, dc);
  const int nrb = ceil_div(wgeom(h).rt, kRTB);
  k_m2l_jobs<<<1, 256, 0, h->stream>>>(G, t->gcount.p, t->m2l_chunk, nrb, R.job_cap, t->gtab.p,
                                       dc);
  k_m2l_fill<<<512, 256, 0, h->stream>>>(G, t->gcount.p, t->m2l_chunk, nrb, gb, t->gtab.p,
                                         (M2LJob *)t->gjobs.p);
  const size_t lds = m2l_lds(wgeom(h));
  k_m2l_mfma_batch<<<t->gm2l_grid, 256, lds, h->stream>>>((const M2LJob *)t->gjobs.p,
                                                  (const M2LArgs *)t->gargs.p, 0, t->gq.p,
                                                  dc + kGcQlen);
  DUST_HIP(hipGetLastError());
  return 0;
}

}  // namespace dust
