// Device-wide primitives for the tree build: stable radix sort of the
// (finest-cell key, particle index) pairs and ordered selection of cells.
#include <hipcub/hipcub.hpp>

#include "fmm_internal.h"

namespace dust {

struct CellPred {
  const unsigned char *flags;
  int leaf_only;
  __host__ __device__ bool operator()(const int &g) const {
    unsigned char f = flags[g];
    return leaf_only ? (f == kLeaf) : (f != 0);
  }
};

// Stable LSD radix sort over the low `key_bits` bits: particles of a cell keep
// their input order, which is the push order of the reference's serial
// binning loop (mod_octree.f90:746-773).
int prim_sort_pairs(FmmHandle *h, int n, int key_bits) {
  size_t bytes = 0;
  DUST_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, h->key.p, h->key_s.p, h->idx_in.p,
                                              h->perm.p, n, 0, key_bits, h->stream));
  DUST_TRY(h->tmp.reserve(bytes));
  DUST_HIP(hipcub::DeviceRadixSort::SortPairs(h->tmp.p, bytes, h->key.p, h->key_s.p, h->idx_in.p,
                                              h->perm.p, n, 0, key_bits, h->stream));
  return 0;
}

// Stable LSD radix sort of (64-bit key, int) pairs over the low `end_bit`
// bits, in the handle's primitive scratch.
int prim_sort_pairs_u64(FmmHandle *h, const unsigned long long *kin, unsigned long long *kout,
                        const int *vin, int *vout, int n, int end_bit) {
  size_t bytes = 0;
  DUST_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, n, 0, end_bit,
                                              h->stream));
  DUST_TRY(h->tmp.reserve(bytes));
  DUST_HIP(hipcub::DeviceRadixSort::SortPairs(h->tmp.p, bytes, kin, kout, vin, vout, n, 0, end_bit,
                                              h->stream));
  return 0;
}

// Global ids of active cells (or of leaves), in ascending order.
int prim_select_flagged_cells(FmmHandle *h, int leaf_only, int *out, int *d_count) {
  hipcub::CountingInputIterator<int> it(0);
  CellPred pred{h->flags.p, leaf_only};
  size_t bytes = 0;
  int nc = (int)h->ncells_tot;
  DUST_HIP(hipcub::DeviceSelect::If(nullptr, bytes, it, out, d_count, nc, pred, h->stream));
  DUST_TRY(h->tmp.reserve(bytes));
  DUST_HIP(hipcub::DeviceSelect::If(h->tmp.p, bytes, it, out, d_count, nc, pred, h->stream));
  return 0;
}

// Sorted positions of the particles a redistribution kept, in order.
int prim_select_kept(FmmHandle *h, const unsigned char *keep, int n, int *out, int *d_count) {
  hipcub::CountingInputIterator<int> it(0);
  size_t bytes = 0;
  DUST_HIP(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, keep, out, d_count, n, h->stream));
  DUST_TRY(h->tmp.reserve(bytes));
  DUST_HIP(hipcub::DeviceSelect::Flagged(h->tmp.p, bytes, it, keep, out, d_count, n, h->stream));
  return 0;
}

}  // namespace dust

namespace dust {

// Size the scratch of the primitives for the handle's capacity once.
int prim_reserve(FmmHandle *h) {
  size_t b1 = 0, b2 = 0;
  int bits = 0;
  while ((1L << bits) < h->nfine) ++bits;
  if (bits == 0) bits = 1;
  DUST_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, h->key.p, h->key_s.p, h->idx_in.p,
                                              h->perm.p, h->cap, 0, bits, h->stream));
  hipcub::CountingInputIterator<int> it(0);
  CellPred pred{h->flags.p, 0};
  DUST_HIP(hipcub::DeviceSelect::If(nullptr, b2, it, h->active.p, h->counts.p,
                                    (int)h->ncells_tot, pred, h->stream));
  return h->tmp.reserve(std::max(b1, b2) + 256);
}

}  // namespace dust
