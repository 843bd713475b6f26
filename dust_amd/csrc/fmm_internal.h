// Internal state of an FMM handle (dust_fmm_t) and the launch interfaces
// shared between the FMM translation units.
#pragma once
#include <vector>

#include "common.h"
#include "wind.h"

namespace dust {

constexpr int kMaxLevels = 10;
constexpr int kSegIters = 1536;  // default P2P source iterations per lane in one item

// Cell flags, the state sort_particles leaves in t_cell%leaf/%branch
// (src/fmm/mod_octree.f90:93-100); active == leaf or branch.
constexpr unsigned char kLeaf = 1;
constexpr unsigned char kBranch = 2;

// Device counts of the graph form of a pass (GraphState::dc): the particles,
// the active cells and leaves, the first active cell of every level, the
// evaluation units and items, the out-of-box particle (atomicMin), the stop
// word (bit 0: particle count outside 0..capacity, 1: active cells or leaves,
// 2: a level, 3: items, 4: M2L jobs over their capacity) and the M2L queue
// length.
enum {
  kGcN = 0,
  kGcActive = 1,
  kGcLeaves = 2,
  kGcActStart = 3,  // kMaxLevels + 1 entries
  kGcUnits = kGcActStart + kMaxLevels + 1,
  kGcItems,
  kGcErr,
  kGcOvf,
  kGcQlen,
  kGcSize = 32
};

// Geometry of one octree level.  Cells of a level are numbered
//   c = box * 8^lvl + morton(ix mod 2^lvl, iy mod 2^lvl, iz mod 2^lvl)
// with box the level-1 box (x fastest) and morton bits interleaved x,y,z from
// the least significant end.  Children of c are 8c+p with p = ic + 2 jc + 4 kc,
// the child order of mod_octree.f90:349-361, and every cell is a contiguous
// range of the finest-level keys, hence of the sorted particles.
struct Level {
  int ncl[3];        // cells per direction
  long ncells;       // cells in the level
  long off;          // first global cell id of the level
  double csize;      // cell edge
  int kdl[3];        // half-extent of the kernel-derivative offset table
  long kd_base;      // first table of the level in the derivative array
};

// Per-stage HIP event pairs for live timing.
struct StageTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, int>> marks;  // (stage, event index pairs)
  double ms[DUST_FMM_NSTAGES] = {0};
  int launches[DUST_FMM_NSTAGES] = {0};
};

// Host side of the small copies between the phases of a pass.
struct FmmReadback {
  int init_err, err;
  int counts[2];                  // active cells, leaves
  int units[2];                   // evaluation units, items
  int act_start[kMaxLevels + 1];  // first active cell of every level
  int ns;                         // particles kept by a redistribution
};

// Graph form of the pass (dust_fmm_graph_*): capacities fixed by
// dust_fmm_graph_reserve, the pass enqueued with every count on the device
// and no host synchronisation, allocation or memset, and optionally captured
// into a graph owned by the handle.
struct GraphState {
  bool reserved = false;
  bool active = false;  // set while the graph form is being enqueued
  int n_cap = 0, act_cap = 0, leaf_cap = 0, unit_cap = 0, item_cap = 0, job_cap = 0;
  int lev_cap[kMaxLevels] = {0};
  DevBuf<int> dc;       // kGcSize device counts
  DevBuf<int> rs_hist;  // radix-sort digit counts (256 per tile) and digit totals
  DevBuf<int> sel_cnt;  // selection counts per tile
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void *> fp;  // buffers the captured graph addresses
};

struct FmmHandle {
  dust_fmm_params prm{};
  int cap = 0, n = 0;
  int nlev = 0;
  Level lev[kMaxLevels];
  long ncells_tot = 0;
  long nfine = 0;
  double xmin[3], xmax[3];
  // polynomial tables
  int deg = 0, nmon = 0;
  int nU = 0;  // derivative entries used by M2L: monomials of order <= 2p
  DevBuf<int> d_pwr;          // (3, nmon)
  DevBuf<double> d_multT;     // (nmon, nmon): multT[n*nmon+m]
  DevBuf<int> d_mnT;          // (nmon, nmon): derivative index of pwr(m)+pwr(n)
  DevBuf<double> d_kd;        // 12*nU per (level, offset): D(3,:) then Dc(3,3,:), SoA
  DevBuf<double> d_fact;      // factorials 0..2p+2
  std::vector<double> h_kd;   // host copy (used by the tuned M2L operator build)
  long kd_tables = 0;
  // particles, input order
  DevBuf<double> cen_in, dir_in, mag_in, rv_in;
  // sort
  DevBuf<unsigned> key, key_s;
  DevBuf<int> idx_in, perm;
  DevBuf<unsigned char> tmp;
  // particles, sorted (SoA)
  DevBuf<double> px, py, pz, dx, dy, dz, pm, pr;
  // tree
  DevBuf<int> fstart;        // nfine + 1
  DevBuf<unsigned char> flags;
  DevBuf<int> slot;
  DevBuf<int> active;        // global ids of active cells, ascending
  DevBuf<int> leaves;        // global ids of leaves, ascending
  DevBuf<int> counts;        // device scratch for list sizes
  int n_active = 0, n_leaves = 0;
  int act_start[kMaxLevels + 1];
  // P2P lists and evaluation work units
  int p2p_cap = 0;
  DevBuf<int2> p2p_rng;
  DevBuf<int> p2p_cnt;
  DevBuf<int> p2p_tot;     // sources in every leaf's P2P list
  DevBuf<int4> units;      // EvalUnit records (2 x int4 each)
  DevBuf<int4> items;      // EvalItem records
  DevBuf<int> n_units;     // device-side unit and item counts
  DevBuf<double> part;     // per item partial P2P sums
  DevBuf<double> leaf_ros;
  int units_cap = 0, n_unit = 0, n_item = 0;
  // expansions
  DevBuf<double> mp_a, mp_b, mp_c;
  // outputs (rot_s / rot: the divergence filter's rotu, with use_divfilt)
  DevBuf<double> vel_s, str_s, tv_s, vel, str, tv, rot_s, rot;
  DevBuf<int> d_err;
  // tuned form state
  void *tuned = nullptr;
  hipStream_t stream = nullptr;
  StageTimer timer;
  int seg_iters = kSegIters;  // P2P source iterations per lane in one item
  bool sorted = false, computed = false, applied = false;
  // set by the batch driver while the tuned M2L and P2P of this pass run in
  // the batch's shared launches (the pass skips its own)
  bool batched = false;
  // page-locked source/targets of the asynchronous copies between pass phases
  HostBuf<FmmReadback> rb;
  // particle redistribution of sort_particles (mod_octree.f90:827-894):
  // settings, element centres, near-wall marks of level pr_lvl (0-based),
  // the kept flags in sorted and input order, the strengths after it in input
  // order, and the staging of the compaction of the sorted arrays
  int pr_on = 0, pr_lvl = 0, pr_m = 0;
  double pr_ratio = 3.0;
  DevBuf<double> pr_cen;
  DevBuf<unsigned char> pr_has, pr_near, pr_keep_s, pr_keep;
  DevBuf<double> pr_dir, pr_mag;
  DevBuf<int> pr_sel, pr_ns, pr_perm;
  DevBuf<unsigned> pr_key;
  DevBuf<double> pr_stage[8];
  bool pr_done = false;  // the last sort ran the redistribution
  // the free-stream term of apply_multipole (variable_wind, time)
  WindDev wind{};
  double wind_time = 0.0;
  int ns = 0;            // particles in the tree after the last sort
  GraphState gr;
};

// fmm_prim.hip: device-wide primitives (radix sort, selection).
int prim_reserve(FmmHandle *h);
int prim_sort_pairs(FmmHandle *h, int n, int key_bits);
int prim_select_flagged_cells(FmmHandle *h, int want_leaf_only, int *out, int *d_count);
int prim_select_kept(FmmHandle *h, const unsigned char *keep, int n, int *out, int *d_count);
int prim_sort_pairs_u64(FmmHandle *h, const unsigned long long *kin, unsigned long long *kout,
                        const int *vin, int *vout, int n, int end_bit);

// fmm_graph.hip: the same primitives with the element count on the device
// (n = *a - (b ? *b : 0), clamped to 0..cap, 0 once *stop is set), sized by
// capacity, with no allocation or memset: the graph form's sort and selection.
struct DCount {
  const int *a, *b, *stop;
  int cap;
};
int gprim_sort_pairs_u32(FmmHandle *h, unsigned *k0, int *v0, unsigned *k1, int *v1, DCount c,
                         int key_bits);
int gprim_sort_pairs_u64(FmmHandle *h, unsigned long long *k0, int *v0, unsigned long long *k1,
                         int *v1, DCount c, int key_bits);
int gprim_select_flagged_cells(FmmHandle *h, int leaf_only, int *out, int cap, int *d_count);
size_t gprim_hist_ints(int cap);
size_t gprim_sel_ints(long n);

// fmm.hip: evaluation helpers shared with the tuned form
struct EvalArgs;
void make_eval_args(const FmmHandle *h, EvalArgs *A);
int eval_leaf_ros(FmmHandle *h, const EvalArgs &A);
int eval_fp64_launch(FmmHandle *h, const EvalArgs &A);
int eval_finish_launch(FmmHandle *h, const EvalArgs &A);

// fmm_tuned.hip
int tuned_prepare(FmmHandle *h);
int tuned_pass_prepare(FmmHandle *h);
int tuned_set(FmmHandle *h, const char *name, double value);
int tuned_get(FmmHandle *h, const char *name, double *value);
int tuned_m2l_all(FmmHandle *h);
int tuned_eval(FmmHandle *h, EvalArgs &A);
int tuned_graph_reserve(FmmHandle *h);
int tuned_graph_m2l(FmmHandle *h);
void tuned_graph_buffers(const FmmHandle *h, std::vector<const void *> &v);
bool tuned_batchable(const FmmHandle *h);
int tuned_batch_launch(FmmHandle *const *hs, int nh, hipStream_t st, hipEvent_t ev_m2l,
                       hipEvent_t ev_p2p);
void tuned_batch_release();
void tuned_release(FmmHandle *h);

// timing helpers
void timer_begin(FmmHandle *h, int stage);
void timer_end(FmmHandle *h, int stage);

}  // namespace dust
