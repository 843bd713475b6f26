// Device-wide primitives of the graph form of the FMM pass: a stable LSD
// radix sort of (key, index) pairs and the ordered selection of flagged
// cells, with the element count read from device memory.  Grids and
// scratch are sized by capacity when the handle is reserved; no call
// allocates, synchronises or issues a memset, so that a stream capture
// records kernel nodes only.  They produce what the hipCUB primitives of the
// host-ordered pass produce: the same stable order (particles of a cell in
// input order, the push order of mod_octree.f90:746-773) and the same
// ascending cell lists.
#include "fmm_internal.h"

namespace dust {

namespace {

constexpr int kTB = 256;          // threads per block
constexpr int kTile = 4 * kTB;    // elements per block: four rounds of 256

__device__ inline int dcount(const DCount &c) {
  if (c.stop && *c.stop) return 0;
  int n = *c.a - (c.b ? *c.b : 0);
  return n < 0 ? 0 : (n > c.cap ? c.cap : n);
}

// Exclusive prefix of one value per thread over the block (256 threads),
// returned with the block total; s is 256 ints of LDS.
__device__ inline int block_excl_scan(int v, int *s, int &total) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int o = 1; o < kTB; o <<= 1) {
    const int a = t >= o ? s[t - o] : 0;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  total = s[kTB - 1];
  const int incl = s[t];
  __syncthreads();
  return incl - v;
}

// Lanes of the wavefront holding the same digit as this lane (valid lanes
// only): one ballot per digit bit.
__device__ inline unsigned long long peers(bool valid, int d, int dbits) {
  unsigned long long m = __ballot(valid);
  for (int b = 0; b < dbits; ++b) {
    const bool bit = (d >> b) & 1;
    const unsigned long long bb = __ballot(valid && bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

// Digit counts of every tile, stored digit-major: hist[d * nb + tile].
template <typename K>
__global__ void __launch_bounds__(kTB)
    k_rs_hist(DCount c, const K *__restrict__ keys, int shift, int dbits, int nb, int *hist) {
  __shared__ int cnt[kTB];
  const int t = threadIdx.x;
  cnt[t] = 0;
  __syncthreads();
  const int n = dcount(c);
  const long base = (long)blockIdx.x * kTile;
  const unsigned mask = (1u << dbits) - 1u;
  for (int r = 0; r < 4; ++r) {
    const long i = base + r * kTB + t;
    if (i < n) atomicAdd(&cnt[(unsigned)(keys[i] >> shift) & mask], 1);
  }
  __syncthreads();
  if (t < (1 << dbits)) hist[(long)t * nb + blockIdx.x] = cnt[t];
}

// Block d: exclusive prefix of digit d's tile counts, in place, and the
// digit total into tot[d].
__global__ void __launch_bounds__(kTB) k_rs_scan(int nb, int *hist, int *tot) {
  __shared__ int s[kTB];
  int *h = hist + (long)blockIdx.x * nb;
  int run = 0;
  for (int b0 = 0; b0 < nb; b0 += kTB) {
    const int q = b0 + (int)threadIdx.x;
    const int v = q < nb ? h[q] : 0;
    int total;
    const int ex = block_excl_scan(v, s, total);
    if (q < nb) h[q] = run + ex;
    run += total;
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = run;
}

// Stable scatter of one digit: position = first slot of the digit (prefix of
// the digit totals) + slots of the digit in earlier tiles + rank in the tile
// (earlier rounds, earlier wavefronts of the round, earlier lanes).
template <typename K>
__global__ void __launch_bounds__(kTB)
    k_rs_scatter(DCount c, const K *__restrict__ kin, const int *__restrict__ vin, K *kout,
                 int *vout, int shift, int dbits, int nb, const int *__restrict__ hist,
                 const int *__restrict__ tot) {
  __shared__ int s[kTB];
  __shared__ int base[kTB];
  __shared__ int wc[4][kTB];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int R = 1 << dbits;
  int total;
  const int dt = t < R ? tot[t] : 0;
  const int ex = block_excl_scan(dt, s, total);
  if (t < R) base[t] = ex + hist[(long)t * nb + blockIdx.x];
  const int n = dcount(c);
  const long b0 = (long)blockIdx.x * kTile;
  const unsigned mask = (1u << dbits) - 1u;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int r = 0; r < 4; ++r) {
    for (int q = t; q < 4 * kTB; q += kTB) (&wc[0][0])[q] = 0;
    __syncthreads();
    const long i = b0 + r * kTB + t;
    const bool valid = i < n;
    K key = 0;
    int d = 0;
    if (valid) {
      key = kin[i];
      d = (int)((unsigned)(key >> shift) & mask);
    }
    const unsigned long long m = peers(valid, d, dbits);
    const int rank = __popcll(m & lt);
    if (valid && rank == 0) wc[w][d] = __popcll(m);
    __syncthreads();
    if (valid) {
      int pos = base[d] + rank;
      for (int v = 0; v < w; ++v) pos += wc[v][d];
      kout[pos] = key;
      vout[pos] = vin[i];
    }
    __syncthreads();
    if (t < R) base[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
    __syncthreads();
  }
}

template <typename K>
int sort_pairs(FmmHandle *h, K *k0, int *v0, K *k1, int *v1, DCount c, int key_bits) {
  // an odd number of passes of <= 8 bits, so the result ends in (k1, v1)
  int passes = (key_bits + 7) / 8;
  if (passes < 1) passes = 1;
  if ((passes & 1) == 0) ++passes;
  const int dbits = (key_bits + passes - 1) / passes;
  const int nb = ceil_div(c.cap, kTile);
  if (nb == 0) return 0;
  DUST_CHECK(h->gr.rs_hist.n >= gprim_hist_ints(c.cap), DUST_E_ARG,
             "graph sort scratch not reserved for %d elements", c.cap);
  int *hist = h->gr.rs_hist.p, *tot = hist + (long)kTB * nb;
  K *src = k0, *dst = k1;
  int *vs = v0, *vd = v1;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * dbits;
    const int db = dbits < key_bits - shift ? dbits : (key_bits - shift > 0 ? key_bits - shift : 1);
    k_rs_hist<K><<<nb, kTB, 0, h->stream>>>(c, src, shift, db, nb, hist);
    k_rs_scan<<<1 << db, kTB, 0, h->stream>>>(nb, hist, tot);
    k_rs_scatter<K><<<nb, kTB, 0, h->stream>>>(c, src, vs, dst, vd, shift, db, nb, hist, tot);
    std::swap(src, dst);
    std::swap(vs, vd);
  }
  DUST_HIP(hipGetLastError());
  return 0;
}

__device__ inline bool cell_pick(const unsigned char *flags, long g, int leaf_only) {
  const unsigned char f = flags[g];
  return leaf_only ? f == kLeaf : f != 0;
}

__global__ void __launch_bounds__(kTB)
    k_sel_count(const unsigned char *__restrict__ flags, long nc, int leaf_only, int *bcnt) {
  __shared__ int s[kTB];
  const long b0 = (long)blockIdx.x * kTile;
  int v = 0;
  for (int r = 0; r < 4; ++r) {
    const long g = b0 + r * kTB + threadIdx.x;
    v += (g < nc && cell_pick(flags, g, leaf_only)) ? 1 : 0;
  }
  int total;
  block_excl_scan(v, s, total);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = total;
}

// One block: exclusive prefix of the tile counts in place, total to *d_count.
__global__ void __launch_bounds__(kTB) k_sel_scan(int nb, int *bcnt, int *d_count) {
  __shared__ int s[kTB];
  int run = 0;
  for (int b0 = 0; b0 < nb; b0 += kTB) {
    const int q = b0 + (int)threadIdx.x;
    const int v = q < nb ? bcnt[q] : 0;
    int total;
    const int ex = block_excl_scan(v, s, total);
    if (q < nb) bcnt[q] = run + ex;
    run += total;
  }
  if (threadIdx.x == 0) *d_count = run;
}

// Selected cells of a tile in ascending order from the tile's first slot;
// slots at or past cap are not written (the count reports the overflow).
__global__ void __launch_bounds__(kTB)
    k_sel_write(const unsigned char *__restrict__ flags, long nc, int leaf_only,
                const int *__restrict__ bcnt, int cap, int *out) {
  __shared__ int wc[4];
  __shared__ int run;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) run = bcnt[blockIdx.x];
  const long b0 = (long)blockIdx.x * kTile;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int r = 0; r < 4; ++r) {
    const long g = b0 + r * kTB + t;
    const bool pick = g < nc && cell_pick(flags, g, leaf_only);
    const unsigned long long m = __ballot(pick);
    if (lane == 0) wc[w] = __popcll(m);
    __syncthreads();
    if (pick) {
      int pos = run + __popcll(m & lt);
      for (int v = 0; v < w; ++v) pos += wc[v];
      if (pos < cap) out[pos] = (int)g;
    }
    __syncthreads();
    if (t == 0) run += wc[0] + wc[1] + wc[2] + wc[3];
    __syncthreads();
  }
}

}  // namespace

// Digit counts of kTB digits per tile plus the kTB digit totals.
size_t gprim_hist_ints(int cap) { return (size_t)kTB * ((size_t)ceil_div(cap, kTile) + 1); }

size_t gprim_sel_ints(long n) { return (size_t)ceil_div(n, (long)kTile) + 1; }

int gprim_sort_pairs_u32(FmmHandle *h, unsigned *k0, int *v0, unsigned *k1, int *v1, DCount c,
                         int key_bits) {
  return sort_pairs<unsigned>(h, k0, v0, k1, v1, c, key_bits);
}

int gprim_sort_pairs_u64(FmmHandle *h, unsigned long long *k0, int *v0, unsigned long long *k1,
                         int *v1, DCount c, int key_bits) {
  return sort_pairs<unsigned long long>(h, k0, v0, k1, v1, c, key_bits);
}

// Global ids of active cells (or of leaves) in ascending order into out
// (at most cap of them), their number into *d_count.
int gprim_select_flagged_cells(FmmHandle *h, int leaf_only, int *out, int cap, int *d_count) {
  const long nc = h->ncells_tot;
  const int nb = (int)ceil_div(nc, (long)kTile);
  DUST_CHECK(h->gr.sel_cnt.n >= gprim_sel_ints(nc), DUST_E_ARG, "graph selection scratch not reserved");
  int *bcnt = h->gr.sel_cnt.p;
  k_sel_count<<<nb, kTB, 0, h->stream>>>(h->flags.p, nc, leaf_only, bcnt);
  k_sel_scan<<<1, kTB, 0, h->stream>>>(nb, bcnt, d_count);
  k_sel_write<<<nb, kTB, 0, h->stream>>>(h->flags.p, nc, leaf_only, bcnt, cap, out);
  DUST_HIP(hipGetLastError());
  return 0;
}

}  // namespace dust
