// Cell numbering, tree geometry and monomial indexing shared by the FMM
// translation units.
#pragma once
#include "fmm_internal.h"

namespace dust {

// ---------------------------------------------------------------------------
// Graded monomial numbering of set_degree_polyexp (mod_multipole.f90:465-488):
// order o = i+j+k, then i descending, then j descending.  Closed form of the
// position: o(o+1)(o+2)/6 monomials of lower order, (o-i)(o-i+1)/2 with a
// larger i, and o-i-j with the same i and a larger j.
// ---------------------------------------------------------------------------
__host__ __device__ inline int mono_index(int i, int j, int k) {
  int o = i + j + k;
  return o * (o + 1) * (o + 2) / 6 + (o - i) * (o - i + 1) / 2 + (o - i - j);
}
inline int mono_count(int deg) { return (deg + 1) * (deg + 2) * (deg + 3) / 6; }

// ---------------------------------------------------------------------------
struct TreeGeom {
  int nlev;
  int nbox[3];
  int ncl[kMaxLevels][3];
  long off[kMaxLevels];
  int kdl[kMaxLevels][3];
  long kd_base[kMaxLevels];
  double xmin[3], xmax[3];
  double csize_fine;
  // Graph form of a pass: the counts live on the device (dc, indexed by
  // kGc*, null in the host-ordered pass) and the kernels size their work
  // from them; the capacities bound what the fixed grids and buffers hold.
  int *dc;
  int act_cap, leaf_cap, item_cap;
};

// The graph form's stop word: a capacity was exceeded or the count is
// invalid, and every later kernel of the pass does nothing.
__device__ inline bool g_stop(const TreeGeom &G) { return G.dc && G.dc[kGcOvf]; }

__host__ __device__ inline unsigned spread3(unsigned v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__host__ __device__ inline unsigned compact3(unsigned v) {
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030C30C3u;
  v = (v | (v >> 4)) & 0x0300F00Fu;
  v = (v | (v >> 8)) & 0x030000FFu;
  v = (v | (v >> 16)) & 0x000003FFu;
  return v;
}

// 0-based cartesian cell coordinates -> cell number within level lvl.
__host__ __device__ inline long cell_encode(const TreeGeom &G, int lvl, int ix, int iy, int iz) {
  int bx = ix >> lvl, by = iy >> lvl, bz = iz >> lvl;
  unsigned msk = (1u << lvl) - 1u;
  long box = bx + (long)G.nbox[0] * (by + (long)G.nbox[1] * bz);
  unsigned m = spread3(ix & msk) | (spread3(iy & msk) << 1) | (spread3(iz & msk) << 2);
  return (box << (3 * lvl)) | (long)m;
}
__host__ __device__ inline void cell_decode(const TreeGeom &G, int lvl, long c, int &ix, int &iy,
                                            int &iz) {
  long box = c >> (3 * lvl);
  unsigned m = (unsigned)(c & ((1L << (3 * lvl)) - 1));
  int bx = (int)(box % G.nbox[0]);
  int by = (int)((box / G.nbox[0]) % G.nbox[1]);
  int bz = (int)(box / ((long)G.nbox[0] * G.nbox[1]));
  ix = (bx << lvl) | (int)compact3(m);
  iy = (by << lvl) | (int)compact3(m >> 1);
  iz = (bz << lvl) | (int)compact3(m >> 2);
}
// Cell centre, evaluated as mod_octree.f90:332-334,363-366.
__host__ __device__ inline double cell_center(const TreeGeom &G, int lvl, int d, int i0) {
  return G.xmin[d] + ((double)(i0 + 1) - 0.5) * (G.xmax[d] - G.xmin[d]) / (double)(G.nbox[d] << lvl);
}
__device__ inline int lvl_of(const TreeGeom &G, long g) {
  int l = 0;
  while (l + 1 < G.nlev && g >= G.off[l + 1]) ++l;
  return l;
}
// Particle range [s, e) of cell c at level lvl in the sorted arrays.
__device__ inline int2 cell_range(const TreeGeom &G, const int *fstart, int lvl, long c) {
  int sh = 3 * (G.nlev - 1 - lvl);
  return make_int2(fstart[c << sh], fstart[(c + 1) << sh]);
}

TreeGeom make_geom(const FmmHandle *h);

}  // namespace dust
