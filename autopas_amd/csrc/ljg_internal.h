// ljg_internal.h -- engine internals shared by the HIP translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ljgpu.h"

namespace ljg {

constexpr int kWave = 64;
constexpr int kMaxTypes = 16;
// Largest tower the LDS bitonic z-sort handles in one workgroup (z: 8 B, id: 8 B, slot: 4 B per entry).
constexpr int kTowerSortCap = 4096;
// Dummy padding slots sit here (far from every box), so that every pair with them fails r^2 <= rc^2.
constexpr double kDummyPos = 1.0e7;
// FP32 cluster-relative position of padding slots; the FP32 force kernel moves padding i-members to
// -kPadPos32 so that padding never meets padding at a finite distance.
constexpr float kPadPos32 = 3.0e30f;

struct HipError {
  hipError_t err;
  const char *what;
  int line;
};

#define LJG_HIP(call)                                         \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) throw ::ljg::HipError{e_, #call, __LINE__}; \
  } while (0)

// Grow-only device buffer.  Growth leaves half again as much room: freeing device memory waits for the
// device, so a rebuild that reallocates stalls the stream (lists lengthen while a lattice melts).
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t cap = 0;
  void ensure(size_t n) {
    if (n <= cap) return;
    if (p) LJG_HIP(hipFree(p));
    p = nullptr;
    size_t c = n + n / 2 + 64;
    LJG_HIP(hipMalloc(&p, c * sizeof(T)));
    cap = c;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Particle state in storage order ("slots").  Slots beyond the real particles of a cluster are padding
// (orig == -1, position kDummyPos, zero velocity/force).
struct State {
  DevBuf<double> x, y, z, vx, vy, vz, fx, fy, fz, ofx, ofy, ofz, rx, ry, rz;
  DevBuf<int32_t> type, orig;
  DevBuf<int64_t> id;
  void ensure(size_t n) {
    for (auto *b : {&x, &y, &z, &vx, &vy, &vz, &fx, &fy, &fz, &ofx, &ofy, &ofz, &rx, &ry, &rz}) b->ensure(n);
    type.ensure(n);
    orig.ensure(n);
    id.ensure(n);
  }
  void release() {
    for (auto *b : {&x, &y, &z, &vx, &vy, &vz, &fx, &fy, &fz, &ofx, &ofy, &ofz, &rx, &ry, &rz}) b->release();
    type.release();
    orig.release();
    id.release();
  }
};

// Plain views of a State for kernel arguments.
struct StateView {
  double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz, *ofx, *ofy, *ofz, *rx, *ry, *rz;
  int32_t *type, *orig;
  int64_t *id;
};

inline StateView view(State &s) {
  return StateView{s.x.p,  s.y.p,  s.z.p,  s.vx.p, s.vy.p, s.vz.p, s.fx.p,   s.fy.p,   s.fz.p,
                   s.ofx.p, s.ofy.p, s.ofz.p, s.rx.p, s.ry.p, s.rz.p, s.type.p, s.orig.p, s.id.p};
}

// CellBlock3D geometry (CellBlock3D.h:359-405) in the form the device needs.
struct CellGeom {
  double box_min[3], box_max[3], recip[3], cell_len[3], L[3];
  int32_t cpil;     // cells per interaction length (1 for csf >= 1)
  int32_t nh[3];    // cells per dimension including halo layers
};

// Periodic tower grid (ClusterTowerBlock2D.h:140-168 without halo towers).
struct TowerGeom {
  double box_min[3], L[3], side[2], recip[2];
  int32_t nt[2];
  int32_t R;        // towers per interaction length
  double ilen;
};

// Tower grid of a box with n particles for cluster size M (ClusterTowerBlock2D::estimateOptimalGridSideLength
// and resize, ClusterTowerBlock2D.h:89-114,140-168, owned towers only).
TowerGeom make_tower_geom(const double lo[3], const double hi[3], int64_t n, int32_t M, double ilen);

// Tower of a position (ClusterTowerBlock2D::getTowerIndex2DAtPosition, ClusterTowerBlock2D.h:220-245).
__device__ inline int32_t tower_index(const TowerGeom &g, double px, double py) {
  const double p[2] = {px, py};
  int32_t t2[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    long t = (long)floor((p[d] - g.box_min[d]) * g.recip[d]);
    if (t < 0) t = 0;
    if (t > g.nt[d] - 1) t = g.nt[d] - 1;
    t2[d] = (int32_t)t;
  }
  return t2[0] + t2[1] * g.nt[0];
}

// 32-bit fixed-point coordinates for the Verlet-list kernels: X = round((x - min) * inv_scale) mod 2^32.
struct FixGeom {
  double min[3], scale[3], inv_scale[3];  // scale = L / 2^32
  float scale32[3];
  int32_t cubic;
};

// A launch that does nothing once an earlier step raised the displacement flag: *flag holds the stamp
// (running step number) of the first drift that moved a particle skin/2 since the last rebuild, and
// stays set until the host has rebuilt.  The step loop queues several steps before it reads the flag;
// every launch of a step at or after the raising one exits at once, so the host resumes from that step
// after the rebuild.  Writes of the flag by the launch itself carry a larger stamp and do not count.
// lst (a j-row build whose list lengths the host has not read yet): the launch also exits when the
// build's statistics show a row longer than its capacity (lst[1] > cap1 for the first pass's sub-rows,
// lst[4] > cap2 for the j rows, lst[5] a sort tile exceeded), or, with frame set, a row that leaves the
// first member's frame (lst[2]).
struct SkipIf {
  const int32_t *flag;
  int32_t stamp;
  const unsigned long long *lst = nullptr;
  uint32_t cap1 = 0, cap2 = 0;
  int32_t frame = 0;
};
__device__ inline bool skip_launch(const SkipIf &s) {
  if (s.lst != nullptr &&
      (s.lst[1] > s.cap1 || s.lst[4] > s.cap2 || s.lst[5] != 0ull || (s.frame != 0 && s.lst[2] != 0ull)))
    return true;
  if (s.flag == nullptr) return false;
  const int32_t f = *s.flag;
  return f != 0 && f <= s.stamp;
}

// Whose potential energy and virial count in the globals.  LJFunctor adds a pair's potential and virial
// to the sums only for an owned i (and an owned j with Newton-3; LJFunctor.h:174-189): with halo
// particles present, a pair of two owned particles counts whole, an owned-halo pair half and a halo-halo
// pair not at all.  The owned particles are those whose upload index is below n; every particle when
// orig is null.
struct Owned {
  const int32_t *orig;  // upload index of each slot (-1: padding)
  int64_t n;
};
__device__ inline bool is_owned(const Owned &o, int64_t s) {
  if (o.orig == nullptr) return true;
  const int32_t k = o.orig[s];
  return k >= 0 && (int64_t)k < o.n;
}

// The index ranges a partial ensemble build works on (the marked boxes' slots, towers or clusters):
// segment j covers [base[j], base[j] + prefix[j+1] - prefix[j]), and a launch over prefix[n] items maps
// item i to its index.  n == 0: the identity over every index (a whole build).
struct Segs {
  const int64_t *prefix, *base;
  int32_t n;
};
__device__ inline int64_t seg_index(const Segs &g, int64_t i) {
  if (g.n == 0) return i;
  int32_t lo = 0, hi = g.n - 1;  // last segment with prefix <= i
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (g.prefix[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return g.base[lo] + (i - g.prefix[lo]);
}

// Velocity-Verlet update fused into a force kernel's epilogue (kick of this step, drift of the next):
// the FP64 state of the current layout, the fixed-point positions to write for the next step, the
// displacement flag (device and pinned host copy) and the per-type factors.
struct IntegArgs {
  double *vx, *vy, *vz, *ofx, *ofy, *ofz, *x, *y, *z;
  const double *rx, *ry, *rz;
  int4 *posfx_next;      // j-row kernel: next fixed-point positions
  float4 *pos32_next;    // FP32 cluster kernel: next cluster-relative positions (relative to c0)
  double4 *posd_next;    // FP32 cluster kernel: next FP64 positions (the copy its cutoff band reads)
  const double4 *c0;
  int32_t *flag, *hflag;
  int32_t *box_flag;            // ensemble: set to 1 for the box whose particle raised the flag (else null)
  const int32_t *cluster_box;   // ensemble: box of each cluster
  double dt, half_skin2;
  int32_t kick, drift, stamp;
  double kick_fac[kMaxTypes];   // dt / (2m)
  double drift_fac[kMaxTypes];  // dt * dt / (2m)
};

// Periodic wrap of positions, as md-flexible's RegularGridDecomposition.cpp:563-587: a coordinate below
// the box moves up by L (at most to the largest value below the upper bound), one at or above the upper
// bound moves down by L (at least to the lower bound).
struct WrapArgs {
  double mn[3], mx[3], L[3], just_inside[3];
};
WrapArgs wrap_args(const ljg_params &p);
WrapArgs wrap_args(const double lo[3], const double hi[3]);
__device__ inline double wrap_coord(double p, int d, const WrapArgs &w) {
  if (p < w.mn[d]) {
    const double q = p + w.L[d];
    return q < w.just_inside[d] ? q : w.just_inside[d];
  }
  if (p >= w.mx[d]) {
    const double q = p - w.L[d];
    return q > w.mn[d] ? q : w.mn[d];
  }
  return p;
}

// Lorentz-Berthelot mixing of per-type epsilon and sigma into ntypes x ntypes tables (row-major):
// ParticlePropertiesLibrary.h:443-473 and calcShift6 :575-582.
inline void mix_table(int32_t ntypes, const double *eps, const double *sig, double rc2, bool apply_shift,
                      double *eps24, double *sig2, double *shift6) {
  for (int a = 0; a < ntypes; ++a) {
    for (int b = 0; b < ntypes; ++b) {
      const int k = ntypes * a + b;
      const double e24 = 24 * std::sqrt(eps[a] * eps[b]);
      const double s = (sig[a] + sig[b]) / 2.0;
      const double s2 = s * s;
      eps24[k] = e24;
      sig2[k] = s2;
      if (apply_shift) {
        const double q2 = s2 / rc2;
        const double q6 = q2 * q2 * q2;
        shift6[k] = e24 * (q6 - q6 * q6);
      } else {
        shift6[k] = 0.;
      }
    }
  }
}

// LJ parameters (mixing tables are ntypes x ntypes, row-major).
struct LJParams {
  int32_t ntypes;
  double rc2;
  double eps24[kMaxTypes * kMaxTypes], sig2[kMaxTypes * kMaxTypes], shift6[kMaxTypes * kMaxTypes];
  double inv2m[kMaxTypes];  // 1/(2m) per type
  double mass[kMaxTypes];
  double eps[kMaxTypes], sig[kMaxTypes];  // per-type parameters the tables were mixed from
};

enum class Layout { None, Upload, Cells, Clusters, Verlet };

// Layout tuning (LJG_KERNEL_AUTO): progress of the current phase and the scores of the last one.
struct TuneState {
  bool active = false;
  int ci = 0;                  // candidate being sampled
  int32_t done = 0;            // its scored steps so far
  bool warmed = false;         // its unscored first step has run
  double wall_ms = 0., rebuild_ms = 0.;
  int32_t rebuilds = 0;
  int best = -1;               // chosen candidate
  int32_t phases = 0;
  int64_t steps_since = 0;     // steps since the last phase ended
  double effective_ms[LJG_MAX_TUNE_CANDIDATES] = {};
};

struct Ctx {
  ljg_params prm{};
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> event_pool;  // timing events of the step loop, reused from call to call
  bool timing = false;

  int64_t n = 0;       // real particles
  int64_t nslots = 0;  // storage slots in the current layout
  // Open box (ljg_set_open_box): the domain [open_min, open_max) and its halo region, embedded in the
  // periodic box prm.box_min/box_max with ilen of empty space beyond the halo, so that no pair meets
  // across the periodic seam.  n_owned: the first n_owned particles of the upload are owned (-1: all).
  bool open = false;
  double open_min[3] = {0., 0., 0.}, open_max[3] = {0., 0., 0.};
  int64_t n_owned = -1;
  Layout layout = Layout::None;
  State st[2];
  int cur = 0;
  LJParams lj{};
  bool types_set = false;

  // LinkedCells
  CellGeom cg{};
  int64_t ncells = 0;
  DevBuf<int32_t> cell_of;      // per slot
  DevBuf<int32_t> cell_count;   // ncells (+1)
  DevBuf<int64_t> cell_start;   // ncells + 1
  DevBuf<int32_t> rank;         // per slot, rank inside cell/tower from the atomic histogram
  DevBuf<int32_t> perm;         // gather map new slot -> old slot
  DevBuf<int64_t> cell_of_particle;  // upload order, for inspection

  // Clusters
  TowerGeom tg{};
  int32_t M = 4;
  int64_t ntowers = 0, nclusters = 0, nentries = 0;
  DevBuf<int32_t> tower_of;     // per old slot
  DevBuf<int32_t> tower_count;  // ntowers
  DevBuf<int64_t> tower_start;  // ntowers + 1 (particle offsets)
  DevBuf<int32_t> tower_items;  // sorted old slots per tower
  DevBuf<double> tower_z;       // z of each tower item
  DevBuf<int64_t> tower_id;     // id of each tower item
  DevBuf<int32_t> ctow_count;   // clusters per tower
  DevBuf<int64_t> ctow_start;   // ntowers + 1 (cluster offsets)
  DevBuf<int32_t> cluster_tower;  // nclusters
  DevBuf<double> bbox;          // 6 * nclusters
  DevBuf<double4> c0;           // nclusters: cluster reference point (bbox centre at build)
  DevBuf<int32_t> list_count;   // entries per i-cluster
  DevBuf<unsigned long long> list_stats;  // list build statistics (totals, longest rows, flags)
  DevBuf<unsigned long long> tstats;      // tower statistics: [0] particles, [1] longest tower, [2] clusters
  int32_t list_row = 0;         // row capacity of the current lists: cluster ci owns [ci*row, ci*row + count)
  int32_t list_cap = 0;         // row capacity to start the next build with
  // List entry: x = j cluster | shift id << 27; y,z,w = FP32 bits of c0_j + shift - c0_i (the offset
  // that turns j's cluster-relative FP32 position into i's frame, formed in FP64 and rounded once).
  DevBuf<int4> list;
  // Member rows of the FP32 cluster kernel: per i-cluster {j slot, entry} of every listed j-cluster
  // member within ilen of one of its members, at stride xstride = list_row * M (no row can outgrow it);
  // xcount per i-cluster.
  DevBuf<int2> xrow, xrow_tmp;
  DevBuf<int32_t> xcount, xcount_tmp;
  int32_t xstride = 0;
  bool xrows_valid = false;     // the current cluster lists carry member rows
  DevBuf<unsigned int> queue;   // work-queue counters of persistent force launches (kQueueRegions + 1)
  bool queue_ready = false;     // the counters were cleared once after allocation
  DevBuf<float4> pos32;         // per slot: position - c0 of its cluster, type bits in w
  DevBuf<float4> pos32_next;    // the fused integrator of the FP32 cluster kernel writes the next step's here
  // FP64 positions of the same step as pos32 (x, y, z; w unused), kept in step with pos32 and swapped with
  // it: the FP32 cluster kernel decides pairs near rc from them, while its fused integrator advances the
  // state's x, y, z in place
  DevBuf<double4> posd, posd_next;
  // bound on |pos32| components since the last full build: kAccSlots lines of kAccStride u64, each the
  // bits of a non-negative float (atomicMax); sets the width of the FP32 cluster kernel's cutoff band
  DevBuf<unsigned long long> pos32_ext;

  // Verlet lists on the linked cells (fixed-point positions, transposed rows per wavefront)
  FixGeom fg{};
  DevBuf<int4> posfx;            // per slot: fixed-point x, y, z and the type
  DevBuf<int4> posfx_next;       // the fused integrator writes the next step's fixed-point positions here
  DevBuf<int32_t> jcl;           // j-row build, first pass: per i-cluster row of j-cluster indices
  int32_t jccap = 0;
  DevBuf<int32_t> jlist;         // cluster layout: per i-cluster row of j slots (union of members' lists)
  DevBuf<int32_t> jcount;        // j slots per i-cluster
  int32_t jcap = 0;
  bool want_jrows = false;       // the next cluster build emits j-particle rows instead of cluster entries
  bool jrows_valid = false;      // the current cluster rows are j-particle rows
  bool jrows_frame_ok = false;   // every row pair within ilen has the same separation in the first member's frame
  // A j-row build the step loop follows with its force at once may leave its list lengths unread
  // (defer_lists, set by the step loop around that build): the force launch then carries a device-side
  // veto on them (SkipIf::lst) and vcl_finish_lists reads them afterwards.
  bool defer_lists = false;
  bool lists_pending = false;    // the current j rows' lengths are not yet checked on the host
  int32_t pend_cap = 0, pend_jcap = 0;
  bool pend_fast = false;        // the vetoed launch used the first member's frame
  bool want_n3 = false;          // the next cluster build emits Newton-3 (forward-only) cluster-entry lists
  bool lists_n3 = false;         // the current cluster-entry lists are Newton-3 lists
  DevBuf<int32_t> vcount;        // neighbours per particle
  DevBuf<int32_t> vlist;         // rows of vrow entries, transposed per wave of 64 particles
  int32_t vrow = 0, vcap = 0;

  DevBuf<double> tab;             // 3 * ntypes^2: (eps24, sig2, shift6) per type pair

  // Reductions
  DevBuf<double> part_e, part_v;  // per work item
  DevBuf<int64_t> part_hits;
  int64_t part_items = 0;         // partial-sum slots written by the last force launch
  unsigned long long *hits_acc = nullptr;  // when set, force kernels add their pair counts here (kAccSlots lines)
  SkipIf skip{nullptr, 0};                 // force launches exit at once when skip_launch(skip)
  DevBuf<unsigned long long> hits_slots;   // the step loop's pair-count lines
  DevBuf<double> red_out;         // [0]=e [1]=v
  DevBuf<int64_t> red_hits;       // [0]
  DevBuf<int32_t> flag;           // [0] displacement, [1] capacity error
  int32_t *host_flag = nullptr;   // pinned, device-mapped: [0] the drift's displacement stamp
  int32_t *host_flag_dev = nullptr;  // device address of host_flag
  unsigned long long *pin_mark = nullptr;      // pinned, device-mapped: [0] marker, [1..] words read back
  unsigned long long *pin_mark_dev = nullptr;  // device address of pin_mark
  unsigned long long mark_seq = 0;

  // Scan scratch
  DevBuf<int64_t> scan_tmp;

  bool have_forces = false;  // st[cur].fx holds the force of the current positions
  int32_t built_M = 0;       // cluster size of the current cluster layout
  TuneState tune;
  int64_t steps_since_rebuild = 0;
  int32_t step_stamp = 0;    // running step number written into flag[0] by the displacement check
  bool carry_force = true;   // permutes keep the current force (false inside the step loop)
  int64_t last_pairs = 0;
  int64_t last_evals = 0;

  // Ensemble of independent boxes (nboxes == 0: the single box of prm).  Upload order is box-major;
  // towers, clusters and slots are too, each box's tower grid being the one it would have alone.
  int32_t nboxes = 0;
  std::vector<double> box_lo, box_hi;  // 3 per box
  std::vector<int64_t> box_count;      // particles per box
  int32_t box_geom_M = 0;              // cluster size the per-box tower grids were made for (0: none yet)
  int64_t box_ntowers = 0;
  DevBuf<TowerGeom> box_tg;            // per box
  DevBuf<WrapArgs> box_wrap;           // per box
  DevBuf<int32_t> box_of;              // box of each particle, upload order
  DevBuf<int64_t> box_tbase;           // nboxes + 1: first tower of each box
  DevBuf<int32_t> tower_box;           // box of each tower
  // Fixed regions: box b owns clusters [cbase[b], cbase[b+1]) (its slots are those clusters' M each),
  // ceil(count/M) + towers of them, enough for any distribution over its towers, so a rebuild of one
  // box moves no other; the clusters after the box's last tower's are empty (all padding, no list).
  // Its tower items (the sort input) are [pbase[b], pbase[b+1]).
  DevBuf<int64_t> box_cbase;           // nboxes + 1
  DevBuf<int64_t> box_pbase;           // nboxes + 1
  DevBuf<int32_t> cluster_box;         // box of each cluster
  int64_t box_clusters = 0;            // cbase[nboxes]
  bool regions_valid = false;          // the slots are laid out in the regions of the current geometry
  bool build_partial = false;          // the next build rebuilds the boxes seg_active marks only, in place
  DevBuf<int32_t> box_flag;            // per box: a particle of the box raised the displacement flag
  DevBuf<int32_t> cell_sorted;         // LinkedCells binning: cell members in id order
  std::vector<int64_t> box_since;      // per box: steps since the box's last rebuild
  DevBuf<int4> list_tmp;               // rows of the boxes a partial build keeps, when the row length grows
  std::vector<int64_t> box_cbase_h, box_tbase_h;  // host copies of box_cbase, box_tbase
  std::vector<int32_t> box_pending;    // per box: placed in its region by an edit, not built yet
  // the marked boxes of a partial build as segments of slots, towers and clusters, and their list
  DevBuf<int64_t> seg_buf;
  std::vector<int64_t> seg_host;
  Segs seg_slots{nullptr, nullptr, 0}, seg_towers{nullptr, nullptr, 0}, seg_clusters{nullptr, nullptr, 0};
  int64_t seg_nslots = 0, seg_ntowers = 0, seg_nclusters = 0;
  const int64_t *seg_boxes = nullptr;
  int32_t seg_nboxes = 0;
  Segs seg_others{nullptr, nullptr, 0};  // the clusters of the boxes not marked
  int64_t seg_nothers = 0;
  int64_t seg_max_count = 0;             // longest tower expected among the marked boxes
  bool partial_tower_sync = false;       // a partial build reads the longest tower back before sorting
  const int32_t *seg_active = nullptr;   // device: the mark per box, uploaded with the segments
  // pinned host staging: the segment upload, and the boxes' displacement flags read back
  char *pin_seg = nullptr;
  size_t pin_seg_cap = 0;
  int32_t *pin_flags = nullptr;
  size_t pin_flags_cap = 0;
  double *pin_dl = nullptr;           // pinned staging of ljg_download_range
  size_t pin_dl_cap = 0;
  // second buffers an edit of the ensemble writes the moved layout into, then swaps in
  DevBuf<float4> pos32_alt, pos32n_alt;
  DevBuf<double4> posd_alt, posdn_alt;
  DevBuf<double> bbox_alt;
  DevBuf<double4> c0_alt;
  DevBuf<int32_t> lcount_alt, ctower_alt;
  // second stream: a partial rebuild and the force of its boxes run beside the force of the others
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // ljg_params.graphs: the executable graph of the last captured chunk of steps (updated or replaced per chunk)
  hipGraphExec_t graph_exec = nullptr;
  int64_t graph_launches = 0, graph_instantiations = 0;
  DevBuf<double> cl_glob;              // per i-cluster (energy x6, virial, pairs) of the last globals launch
  DevBuf<double> box_glob;             // per box (energy, virial, pairs) after endTraversal scaling
  bool box_glob_valid = false;
};

// ---- launchers (defined in the .hip translation units) -------------------------------------------
void exclusive_scan_i32_to_i64(Ctx &c, const int32_t *in, int64_t *out, int64_t n);  // out has n+1 entries

void lc_bin(Ctx &c);                       // ljg_bin.hip
// CellBlock3D geometry of the box [lo, hi) for interaction length ilen, cell size factor 1 (ljg_bin.hip)
CellGeom make_cell_geom(const double lo[3], const double hi[3], double ilen, int64_t *ncells);
// LinkedCells binning of n particles on explicit buffers (ljg_bin.hip): phase 0 = cell and rank of every
// particle (count must be zero), phase 1 = scatter into the cells by rank and order each cell by id (index
// when id is null) into sorted.
void cell_bin_launch(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const CellGeom &g,
                     int32_t *cell_of, int32_t *rank, int32_t *count, const int64_t *start, int32_t *items,
                     const int64_t *id, int32_t *sorted, int phase);
// The plain FP64 linked-cells kernel on explicit buffers (ljg_force.hip); block partials into pe/pv/ph
// (blocks_for(n, 256) of them).
void lc_force_launch(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                     const int32_t *type, const int32_t *cell_of, const int64_t *cell_start, const CellGeom &g,
                     double rc2, int32_t ntypes, const double *tab, double eps24, double sig2, double shift6,
                     double *fx, double *fy, double *fz, double *pe, double *pv, int64_t *ph, bool globals,
                     unsigned long long *acc, SkipIf skip, Owned owned = Owned{nullptr, 0});
// Exclusive scan on a given stream with caller-provided temporary storage of scan_tmp_len(n) entries.
int64_t scan_tmp_len(int64_t n);  // ljg_md.hip
void exclusive_scan_i32_to_i64_on(hipStream_t st, const int32_t *in, int64_t *out, int64_t n, int64_t *tmp,
                                  int64_t tmp_len);
// Deterministic sum of block partials into out[0] (energy x6), out[1] (virial), out_h[0] (pair hits).
void reduce_launch(hipStream_t st, int64_t nitems, const double *pe, const double *pv, const int64_t *ph,
                   double *out, int64_t *out_h, bool globals);  // ljg_md.hip
void lc_force(Ctx &c, bool globals);       // ljg_bin.hip (plain FP64 kernel)
void vcl_build(Ctx &c);                    // ljg_vcl.hip
// Reads and checks the list lengths of a j-row build that left them pending (true when none is pending):
// false when a row overflowed or the vetoed force launch assumed a frame the rows leave, in which case
// that launch exited without effect and the caller builds again and launches again.  ljg_vcl.hip
bool vcl_finish_lists(Ctx &c);
void fill_i32(Ctx &c, int32_t *p, int64_t n, int32_t v);  // ljg_vcl.hip: p[0..n) = v on the stream
void vcl_pos32(Ctx &c, const int32_t *active = nullptr);  // ljg_vcl.hip (active: ensemble boxes to write)
// A launch of the FP32 cluster kernel over some i-clusters only (the segments' nitems clusters), with
// its block partials from part_offset (reserved by the caller).
// Work queue of a persistent force launch: the tiles (G i-clusters of one wavefront) are taken in chunks
// of `chunk` consecutive tiles from kQueueRegions counters, one per contiguous region of the tiles, each
// wavefront starting at its XCD's region (workgroup b runs on XCD b % 8) and moving on to the next
// regions when its own is used up.  ctr[kQueueRegions] counts the wavefronts that are done; the last one
// clears every counter, so the next launch finds them at zero.  ctr null: one tile per wavefront, given
// by the launch grid.
constexpr int kQueueRegions = 8;
struct WorkQueue {
  unsigned int *ctr;
  int32_t ntiles, chunk, nwaves;
};

struct ClusterSubset {
  Segs segs;
  int64_t nitems;
  int64_t part_offset;
};
// n3: the lists are Newton-3 lists and each pair's force goes to both particles (the force arrays are
// cleared first and every contribution is added atomically).  ljg_force.hip
void cluster_force(Ctx &c, bool mixed, bool globals, const IntegArgs *integ = nullptr,
                   const ClusterSubset *sub = nullptr, bool n3 = false);
int64_t cluster_force_blocks(const Ctx &c, int64_t nitems);  // ljg_force.hip: workgroups of a launch
void verlet_build(Ctx &c);                 // ljg_verlet.hip (needs the cell layout)
void verlet_fixed(Ctx &c);                 // ljg_verlet.hip
void verlet_force(Ctx &c, bool globals);   // ljg_verlet.hip
void verlet_geometry(Ctx &c);              // ljg_verlet.hip: fixed-point scales of the box
void verlet_fixed_slots(Ctx &c, int64_t nslots);  // ljg_verlet.hip: posfx for slots [0, nslots)
void cluster_jrows_force(Ctx &c, bool globals, const IntegArgs *integ = nullptr);  // ljg_force.hip
IntegArgs integ_args(Ctx &c, bool kick, bool drift, double dt, int32_t stamp);    // ljg_md.hip
void md_kick_drift(Ctx &c, bool kick, bool drift, double dt, int32_t stamp, SkipIf skip = SkipIf{nullptr, 0});  // ljg_md.hip
void wrap_positions(Ctx &c);               // ljg_md.hip
// Rebuild permutation (also sets the rebuild reference positions; writes posfx when given).
void permute_state(Ctx &c, const int32_t *src, int64_t new_slots, int4 *posfx = nullptr);  // ljg_md.hip
void reduce_partials(Ctx &c, int64_t nitems, ljg_globals *out, bool globals);  // ljg_md.hip
// Positions (3n: x, y, z in upload order) into the current layout, velocities too when vel != nullptr,
// the derived FP32 / fixed-point positions refreshed; flag[2] = 1 if a particle moved skin/2 since the
// last rebuild.  ljg_md.hip
void md_set_positions(Ctx &c, const double *pos, const double *vel);
// Waits until the stream has run everything queued, and reads n (<= 15) words of device memory back:
// a one-thread kernel copies them into pinned mapped memory and then writes a fresh marker there, on
// which the host spins (a stream synchronisation wakes the host tens of microseconds after the work
// ends).  After 2 s of spinning it synchronises the stream instead.  ljg_md.hip
constexpr int kMarkWords = 63;                          // words one marker read returns at most
constexpr int kMarkFlagBoxes = 2 * (kMarkWords - 1);    // ensembles up to this size read box flags with it
// Waits for the work queued on the stream and reads n words of src, then n2 words of src2 (device
// memory), into out, with one marker kernel whose pinned word the host polls.
void stream_wait(Ctx &c, const unsigned long long *src = nullptr, int n = 0, unsigned long long *out = nullptr,
                 const unsigned long long *src2 = nullptr, int n2 = 0);
// The two halves of stream_wait, for work that can be queued between them (it runs while the host waits
// for the marker): stream_mark queues the marker that copies n words of src and returns its sequence
// number; stream_spin waits for that marker and reads the n words into out.  No other marker may be
// queued in between.
unsigned long long stream_mark(Ctx &c, const unsigned long long *src, int n);
void stream_spin(Ctx &c, unsigned long long seq, int n, unsigned long long *out);
// stats[0] += sum(cnt[0..n)), stats[1] = max(stats[1], max(cnt)); a few hundred atomics at most.
void count_stats(Ctx &c, const int32_t *cnt, int64_t n, unsigned long long *stats);  // ljg_md.hip
// Ensemble (ljg_ensemble.hip): per-box tower grids and fixed regions for cluster size c.M, the tower
// histogram that wraps each position into its own box, the per-box tower layout inside the regions, the
// wrap alone, and the per-box reduction of cl_glob into box_glob.  `active` (device, per box; null: every
// box) restricts a build to the boxes it marks.
void ensemble_geometry(Ctx &c);
// owner[i] = the b with start[b] <= i < start[b + 1] for i in [0, n) (start: nb + 1 offsets, on the device)
void fill_owner(Ctx &c, int32_t *owner, int64_t n, const int64_t *start, int32_t nb);
void ensemble_tower_hist(Ctx &c, const int32_t *active);
void ensemble_layout(Ctx &c, const int32_t *active);
// src = -1 and cluster_tower = -1 over the regions of the active boxes (the sorts then fill their towers)
void ensemble_region_prep(Ctx &c, const int32_t *active);
// The cluster-pair rows of the boxes that `active` does not mark, from stride oldcap to newcap.
void ensemble_restride_lists(Ctx &c, const int32_t *active, int32_t oldcap, int32_t newcap);
void ensemble_wrap(Ctx &c);
void ensemble_box_globals(Ctx &c);
// An edit of an ensemble laid out in its regions: the tables of the new box list are set (kept boxes
// first, in keep order, then the added ones) and each kept box's slots, clusters, cluster-pair rows and
// FP32 positions move to its new region unchanged (cluster indices in the rows shifted with it); the
// added particles are placed in their boxes' regions, unsorted, for a build of those boxes
// (box_pending).  src_box[k] is the old index of kept box k; old_pstart the old boxes' first upload
// indices.
void ensemble_edit_regions(Ctx &c, const std::vector<int32_t> &src_box, const std::vector<int64_t> &old_pstart,
                           int64_t kept, int64_t added, const double *x, const double *y, const double *z,
                           const double *vx, const double *vy, const double *vz, const int32_t *type,
                           const int64_t *id);
// The slots of the active boxes' regions gathered from src (old slots of the same regions) in place,
// through the other state buffer; the other boxes' slots are not touched.  Nothing moves when *skip is
// set on the device (skip may be null).  ljg_md.hip
void permute_active(Ctx &c, const int32_t *src, const int32_t *active, const unsigned long long *skip = nullptr);
// The member rows of the FP32 cluster kernel after a cluster-entry build (all clusters, or with `partial`
// the marked boxes' of an ensemble).
void vcl_member_rows(Ctx &c, bool partial, const int32_t *active);  // ljg_vcl.hip
// Keep the boxes with new_base[b] >= 0 (their particles move to upload indices new_base[b] + local index,
// with positions, velocities, forces, types and ids), then append `added` new particles from the host
// after the `kept` ones (zero force).  Leaves the state in upload order in the other buffer, made current.
void ensemble_edit(Ctx &c, const std::vector<int64_t> &old_pstart, const std::vector<int64_t> &new_base, int64_t kept,
                   int64_t added, const double *x, const double *y, const double *z, const double *vx,
                   const double *vy, const double *vz, const int32_t *type, const int64_t *id);

constexpr int kJclSplit = 4;   // wavefronts per tower in the first pass of the j-row build (list_count holds
                                // kJclSplit counts per cluster)
constexpr int kShiftBits = 27;
constexpr int32_t kClusterMask = (1 << kShiftBits) - 1;

// Shift id (sx+1) + 3(sy+1) + 9(sz+1), sx,sy,sz in {-1,0,1}.
__host__ __device__ inline int shift_component(int sid, int d) {
  return (d == 0 ? sid % 3 : d == 1 ? (sid / 3) % 3 : sid / 9) - 1;
}

inline int blocks_for(int64_t n, int threads) { return (int)((n + threads - 1) / threads); }

// The FP32 cluster kernel's two position buffers change roles (its fused integrator wrote the next step's
// positions into the second): the FP32 and the FP64 copies always together.
// The ownership the force kernels mask the globals with: none (every particle counts) unless the caller
// marked fewer owned particles than there are.
inline Owned owned_of(Ctx &c) {
  if (c.n_owned < 0 || c.n_owned >= c.n) return Owned{nullptr, 0};
  return Owned{c.st[c.cur].orig.p, c.n_owned};
}

inline void swap_cluster_positions(Ctx &c) {
  std::swap(c.pos32, c.pos32_next);
  std::swap(c.posd, c.posd_next);
}

// Thresholds on bits(r^2) - 1 (r^2 a positive FP32) around rc2 for the FP32-pair kernels that decide
// the pairs near the cutoff again in FP64: below lo the pair is inside, from hi on outside.  rel is the
// relative half-width of the band (1 ulp of rc2 is 2^-23 to 2^-24 of it).
inline void cutoff_band(float rc2, double rel, uint32_t &lo, uint32_t &hi) {
  uint32_t b;
  std::memcpy(&b, &rc2, sizeof b);
  const double ulps = std::ceil(rel * 8388608.0);
  const uint32_t w = (uint32_t)(ulps < 1.0e6 ? ulps : 1.0e6);
  lo = b - w;
  hi = b + w;
}

// Same-address device atomics from many workgroups serialize (~12 ns each on MI355X); counters that
// every workgroup of a kernel updates are spread over kAccSlots lines, one per XCD (workgroups are
// dealt to the 8 XCDs round-robin), and summed by the host.
constexpr int kAccSlots = 8;
constexpr int kAccStride = 16;  // u64 per slot: one 128-byte line each
__device__ inline void acc_add(unsigned long long *acc, unsigned long long v) {
  atomicAdd(acc + (blockIdx.x & (kAccSlots - 1)) * kAccStride, v);
}

// lines[l * kAccStride] = max(lines[...], v) over the workgroup's v >= 0 (as float bits, so the unsigned
// order is the float order), with one atomic per workgroup at most, on the line of its XCD, and none when
// the line already holds as much.  Every thread of the workgroup (blockDim.x a multiple of 64, at most
// 1024) must call this.
__device__ inline void block_max_to_lines(float v, unsigned long long *lines) {
  __shared__ float wmax[16];
  for (int off = kWave / 2; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
  const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) wmax[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    for (int k = 0; k < nw; ++k) m = fmaxf(m, wmax[k]);
    unsigned long long *line = lines + (blockIdx.x & (kAccSlots - 1)) * kAccStride;
    const unsigned long long bits = (unsigned long long)__float_as_uint(m);
    if (bits > *(volatile unsigned long long *)line) atomicMax(line, bits);
  }
}

// Histogram rank with one atomic per run of equal keys inside the wavefront.  Particles arrive in the
// previous build's order, so neighbouring lanes mostly share a cell/tower and the atomics drop ~30x.
// Every lane of the wave must call this (inactive lanes pass active = false); returns the slot inside
// the bin for active lanes.
__device__ inline int32_t wave_aggregated_rank(int32_t key, bool active, int32_t *count) {
  const int lane = threadIdx.x & (kWave - 1);
  const int32_t k = active ? key : -1;
  const int32_t prev = __shfl_up(k, 1, kWave);
  const bool head = lane == 0 || k != prev;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
  const int leader = 63 - __clzll((long long)(heads & upto));
  const unsigned long long after = heads & ~upto;
  const int next = after ? __ffsll((long long)after) - 1 : kWave;
  int32_t base = 0;
  if (lane == leader && k >= 0) base = atomicAdd(&count[k], next - leader);
  base = __shfl(base, leader, kWave);
  return base + (lane - leader);
}

}  // namespace ljg
