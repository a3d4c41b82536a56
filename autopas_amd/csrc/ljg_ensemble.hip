// ljg_ensemble.hip -- many independent periodic boxes in one context.
//
// md-flexible runs one AutoPas instance per simulation (AutoPasImpl.h:55-113), each with its own tower
// grid (ClusterTowerBlock2D.h:140-168), cluster lists and rebuilds.  Here the boxes' particles are
// concatenated box by box and the force launch covers all boxes:
//   * each box gets the tower grid it would have alone (make_tower_geom on its own box and count), and
//     the boxes' towers are numbered one box after the other, so clusters and slots stay box-major;
//   * each box owns a fixed region of clusters (and their slots), ceil(count/M) + towers of them, which
//     holds any distribution of its particles over its towers; the clusters after its last tower's are
//     empty.  A box is rebuilt in place in its region, so one box's rebuild moves no other box, and
//     every box keeps the rebuild schedule it has alone (the step loop's per-box flags and counters);
//   * the tower histogram wraps each particle into its own box (RegularGridDecomposition.cpp:563-587);
//   * the cluster-pair list pass reads the grid of the tower's box (k_cluster_lists_tower<.., MB>), and
//     its list entries carry the image offset c0_j + shift - c0_i formed with that box's lengths, so the
//     FP32 cluster force kernel runs on the ensemble unchanged;
//   * an edit of the box list moves each kept box's region (slots, clusters, rows, FP32 positions) as
//     it is, and places the added boxes' particles in their regions for a build of those boxes alone;
//   * per-box potential energy, virial and pair count are reduced from per-cluster sums in a fixed order.
// Compiled with -ffp-contract=off like the single-box build: a box's grid, wrap and tower indices round
// exactly as they do when the box is alone.
#include <cmath>

#include "ljg_internal.h"

namespace ljg {

// Tower histogram over the ensemble: wrap into the particle's own box, then that box's tower index
// offset by the box's first tower.
__global__ void k_tower_hist_boxes(int64_t nitems, Segs sg, double *__restrict__ x, double *__restrict__ y,
                                   double *__restrict__ z, const int32_t *__restrict__ orig,
                                   const int32_t *__restrict__ box_of, const TowerGeom *__restrict__ tg,
                                   const WrapArgs *__restrict__ wr, const int64_t *__restrict__ tbase,
                                   const int32_t *__restrict__ box_active, int32_t *__restrict__ tower_of,
                                   int32_t *__restrict__ rank, int32_t *__restrict__ count) {
  const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = it < nitems ? seg_index(sg, it) : -1;
  const int32_t o = i >= 0 ? orig[i] : -1;
  const int32_t b = o >= 0 ? box_of[o] : 0;
  const bool active = o >= 0 && (box_active == nullptr || box_active[b] != 0);
  int32_t t = -1;
  if (active) {
    const WrapArgs w = wr[b];
    const double ox = x[i], oy = y[i], oz = z[i];
    const double px = wrap_coord(ox, 0, w), py = wrap_coord(oy, 1, w), pz = wrap_coord(oz, 2, w);
    if (px != ox) x[i] = px;
    if (py != oy) y[i] = py;
    if (pz != oz) z[i] = pz;
    t = (int32_t)(tbase[b] + tower_index(tg[b], px, py));
  }
  const int32_t r = wave_aggregated_rank(t, active, count);
  if (i >= 0) {
    tower_of[i] = t;
    rank[i] = r;
  }
}

__global__ void k_wrap_boxes(int64_t nslots, StateView s, const int32_t *__restrict__ box_of,
                             const WrapArgs *__restrict__ wr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots || s.orig[i] < 0) return;
  const WrapArgs w = wr[box_of[s.orig[i]]];
  double *c[3] = {s.x, s.y, s.z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double p = c[d][i], q = wrap_coord(p, d, w);
    if (q != p) c[d][i] = q;
  }
}

// Tower layout of each box inside its regions, one wavefront per box (marked boxes only when box_active
// is given): particle offsets start[t] from pbase[b], cluster offsets cstart[t] from cbase[b] with
// ccount[t] = ceil(count[t] / M) clusters, towers in index order; stats[1] = longest tower.
constexpr int kBoxLayoutThreads = 256;

__global__ __launch_bounds__(kBoxLayoutThreads) void k_box_layout(
    int32_t nboxes, const int64_t *__restrict__ boxes, const int64_t *__restrict__ tbase,
    const int64_t *__restrict__ pbase, const int64_t *__restrict__ cbase, const int32_t *__restrict__ box_active,
    const int32_t *__restrict__ count, int32_t M, int64_t *__restrict__ start, int32_t *__restrict__ ccount,
    int64_t *__restrict__ cstart, unsigned long long *__restrict__ stats) {
  const int lane = threadIdx.x & (kWave - 1);
  const int32_t w = (int32_t)(blockIdx.x * (kBoxLayoutThreads / kWave) + threadIdx.x / kWave);
  if (w >= nboxes) return;
  const int32_t b = boxes ? (int32_t)boxes[w] : w;  // the marked boxes' list, or every box
  if (box_active && box_active[b] == 0) return;
  const int64_t t0 = tbase[b], t1 = tbase[b + 1];
  int64_t p = pbase[b], q = cbase[b];
  int32_t mx = 0;
  for (int64_t tb = t0; tb < t1; tb += kWave) {
    const int64_t t = tb + lane;
    const int32_t k = t < t1 ? count[t] : 0;
    const int32_t kc = (k + M - 1) / M;
    int32_t sk = k, sc = kc;  // inclusive scans over the lanes
    for (int off = 1; off < kWave; off <<= 1) {
      const int32_t ak = __shfl_up(sk, off, kWave), ac = __shfl_up(sc, off, kWave);
      if (lane >= off) {
        sk += ak;
        sc += ac;
      }
    }
    if (t < t1) {
      start[t] = p + (sk - k);
      cstart[t] = q + (sc - kc);
      ccount[t] = kc;
    }
    p += __shfl(sk, kWave - 1, kWave);
    q += __shfl(sc, kWave - 1, kWave);
    mx = k > mx ? k : mx;
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const int32_t o = __shfl_xor(mx, off, kWave);
    mx = o > mx ? o : mx;
  }
  if (lane == 0) atomicMax(stats + 1, (unsigned long long)mx);
}

// src = -1 (padding) over the slots of the marked boxes' regions, and no tower for their clusters.
__global__ void k_region_prep(int64_t nitems, Segs sg, int32_t M, const int32_t *__restrict__ cluster_box,
                              const int32_t *__restrict__ box_active, int32_t *__restrict__ src,
                              int32_t *__restrict__ cluster_tower) {
  const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nitems) return;
  const int64_t s = seg_index(sg, it);
  const int64_t c = s / M;
  if (box_active && box_active[cluster_box[c]] == 0) return;
  src[s] = -1;
  if (s % M == 0) cluster_tower[c] = -1;
}

// A wavefront per cluster of an unmarked box copies its row (count[c] <= oldcap entries) to stride newcap.
constexpr int kRestrideThreads = 256;

__global__ __launch_bounds__(kRestrideThreads) void k_restride_rows(
    int64_t nclusters, const int32_t *__restrict__ cluster_box, const int32_t *__restrict__ box_active,
    const int32_t *__restrict__ count, const int4 *__restrict__ in, int32_t oldcap, int4 *__restrict__ out,
    int32_t newcap) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t c = (int64_t)blockIdx.x * (kRestrideThreads / kWave) + threadIdx.x / kWave;
  if (c >= nclusters || box_active[cluster_box[c]] != 0) return;
  const int32_t n = count[c] < oldcap ? count[c] : oldcap;
  for (int32_t k = lane; k < n; k += kWave) out[c * (int64_t)newcap + k] = in[c * (int64_t)oldcap + k];
}

// One wavefront per box: the sums of the box's i-clusters [cbase[b], cbase[b+1]) (the empty ones add 0)
// in a fixed lane assignment and butterfly, then LJFunctor::endTraversal's scaling for full lists (halve
// everything, energy / 6; LJFunctor.h:661-685).
constexpr int kBoxReduceThreads = 256;

__global__ __launch_bounds__(kBoxReduceThreads) void k_box_reduce(int32_t nboxes, const int64_t *__restrict__ cbase,
                                                                  const double *__restrict__ cl,
                                                                  double *__restrict__ out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int32_t b = (int32_t)(blockIdx.x * (kBoxReduceThreads / kWave) + threadIdx.x / kWave);
  if (b >= nboxes) return;
  const int64_t cb = cbase[b], ce = cbase[b + 1];
  double e = 0., v = 0., h = 0.;
  for (int64_t ci = cb + lane; ci < ce; ci += kWave) {
    e += cl[3 * ci];
    v += cl[3 * ci + 1];
    h += cl[3 * ci + 2];
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, kWave);
    v += __shfl_xor(v, off, kWave);
    h += __shfl_xor(h, off, kWave);
  }
  if (lane == 0) {
    out[3 * b] = e * 0.5 / 6.;
    out[3 * b + 1] = v * 0.5;
    out[3 * b + 2] = h * 0.5;
  }
}

// Particles of kept boxes from the current layout to their new upload index (the rebuild reference is
// reset to the position: the edit is followed by a rebuild).
__global__ void k_keep_boxes(int64_t nslots, StateView o, StateView n, const int32_t *__restrict__ box_of,
                             const int64_t *__restrict__ old_pstart, const int64_t *__restrict__ new_base) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const int32_t orig = o.orig[s];
  if (orig < 0) return;
  const int32_t b = box_of[orig];
  const int64_t nb = new_base[b];
  if (nb < 0) return;
  const int64_t i = nb + (orig - old_pstart[b]);
  const double x = o.x[s], y = o.y[s], z = o.z[s];
  n.x[i] = x; n.y[i] = y; n.z[i] = z;
  n.rx[i] = x; n.ry[i] = y; n.rz[i] = z;
  n.vx[i] = o.vx[s]; n.vy[i] = o.vy[s]; n.vz[i] = o.vz[s];
  n.fx[i] = o.fx[s]; n.fy[i] = o.fy[s]; n.fz[i] = o.fz[s];
  n.ofx[i] = o.ofx[s]; n.ofy[i] = o.ofy[s]; n.ofz[i] = o.ofz[s];
  n.type[i] = o.type[s];
  n.orig[i] = (int32_t)i;
  n.id[i] = o.id[s];
}

// New particles (6 x count doubles: x, y, z, vx, vy, vz) at upload indices first .. first + count.
__global__ void k_add_particles(int64_t first, int64_t count, StateView n, const double *__restrict__ in,
                                int32_t have_v, const int32_t *__restrict__ type, const int64_t *__restrict__ id) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int64_t i = first + k;
  const double x = in[k], y = in[count + k], z = in[2 * count + k];
  n.x[i] = x; n.y[i] = y; n.z[i] = z;
  n.rx[i] = x; n.ry[i] = y; n.rz[i] = z;
  n.vx[i] = have_v ? in[3 * count + k] : 0.;
  n.vy[i] = have_v ? in[4 * count + k] : 0.;
  n.vz[i] = have_v ? in[5 * count + k] : 0.;
  n.fx[i] = 0.; n.fy[i] = 0.; n.fz[i] = 0.;
  n.ofx[i] = 0.; n.ofy[i] = 0.; n.ofz[i] = 0.;
  n.type[i] = type ? type[k] : 0;
  n.orig[i] = (int32_t)i;
  n.id[i] = id ? id[k] : i;
}

void ensemble_edit(Ctx &c, const std::vector<int64_t> &old_pstart, const std::vector<int64_t> &new_base, int64_t kept,
                   int64_t added, const double *x, const double *y, const double *z, const double *vx,
                   const double *vy, const double *vz, const int32_t *type, const int64_t *id) {
  const int64_t total = kept + added;
  const int other = 1 - c.cur;
  c.st[other].ensure((size_t)total);
  if (kept > 0) {
    c.scan_tmp.ensure(old_pstart.size() + new_base.size());
    int64_t *d_pstart = c.scan_tmp.p, *d_base = c.scan_tmp.p + old_pstart.size();
    LJG_HIP(hipMemcpyAsync(d_pstart, old_pstart.data(), old_pstart.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           c.stream));
    LJG_HIP(hipMemcpyAsync(d_base, new_base.data(), new_base.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           c.stream));
    hipLaunchKernelGGL(k_keep_boxes, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots,
                       view(c.st[c.cur]), view(c.st[other]), c.box_of.p, d_pstart, d_base);
    LJG_HIP(hipGetLastError());
    LJG_HIP(hipStreamSynchronize(c.stream));  // the host tables go out of scope
  }
  if (added > 0) {
    const bool have_v = vx && vy && vz;
    const size_t n = (size_t)added;
    c.part_v.ensure(6 * n);
    const double *src[6] = {x, y, z, vx, vy, vz};
    for (int k = 0; k < (have_v ? 6 : 3); ++k)
      LJG_HIP(hipMemcpyAsync(c.part_v.p + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    int32_t *d_type = nullptr;
    int64_t *d_id = nullptr;
    if (type) {
      c.rank.ensure(n);
      d_type = c.rank.p;
      LJG_HIP(hipMemcpyAsync(d_type, type, n * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    }
    if (id) {
      c.scan_tmp.ensure(n);
      d_id = c.scan_tmp.p;
      LJG_HIP(hipMemcpyAsync(d_id, id, n * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
    }
    hipLaunchKernelGGL(k_add_particles, dim3(blocks_for(added, 256)), dim3(256), 0, c.stream, kept, added,
                       view(c.st[other]), c.part_v.p, (int32_t)have_v, d_type, d_id);
    LJG_HIP(hipGetLastError());
    LJG_HIP(hipStreamSynchronize(c.stream));
  }
  c.cur = other;
}

// ---- edits that keep the running boxes' layouts ----------------------------------------------------
// New slot s of new box k = cluster_box[s / M] comes from old slot s + (src_cbase[k] - cbase[k]) M when
// the box is kept (upload index shifted by orig_delta[k]); the slots of added boxes start as padding.
__global__ void k_move_slots(int64_t nslots, int32_t M, const int32_t *__restrict__ cluster_box,
                             const int64_t *__restrict__ cbase, const int64_t *__restrict__ src_cbase,
                             const int64_t *__restrict__ orig_delta, StateView o, StateView n,
                             const float4 *__restrict__ pos32_old, float4 *__restrict__ pos32,
                             float4 *__restrict__ pos32_next, const double4 *__restrict__ posd_old,
                             double4 *__restrict__ posd, double4 *__restrict__ posd_next) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const int32_t k = cluster_box[s / M];
  const int64_t sc = src_cbase[k];
  if (sc < 0) {
    n.x[s] = kDummyPos; n.y[s] = kDummyPos; n.z[s] = kDummyPos;
    n.vx[s] = 0.; n.vy[s] = 0.; n.vz[s] = 0.;
    n.fx[s] = 0.; n.fy[s] = 0.; n.fz[s] = 0.;
    n.ofx[s] = 0.; n.ofy[s] = 0.; n.ofz[s] = 0.;
    n.rx[s] = kDummyPos; n.ry[s] = kDummyPos; n.rz[s] = kDummyPos;
    n.type[s] = 0; n.orig[s] = -1; n.id[s] = -1;
    const float4 p = make_float4(kPadPos32, kPadPos32, kPadPos32, __int_as_float(0));
    pos32[s] = p;
    pos32_next[s] = p;
    const double4 d = make_double4(kDummyPos, kDummyPos, kDummyPos, 0.);
    posd[s] = d;
    posd_next[s] = d;
    return;
  }
  const int64_t f = s + (sc - cbase[k]) * M;
  n.x[s] = o.x[f];   n.y[s] = o.y[f];   n.z[s] = o.z[f];
  n.vx[s] = o.vx[f]; n.vy[s] = o.vy[f]; n.vz[s] = o.vz[f];
  n.fx[s] = o.fx[f]; n.fy[s] = o.fy[f]; n.fz[s] = o.fz[f];
  n.ofx[s] = o.ofx[f]; n.ofy[s] = o.ofy[f]; n.ofz[s] = o.ofz[f];
  n.rx[s] = o.rx[f]; n.ry[s] = o.ry[f]; n.rz[s] = o.rz[f];
  n.type[s] = o.type[f];
  const int32_t og = o.orig[f];
  n.orig[s] = og >= 0 ? (int32_t)(og + orig_delta[k]) : -1;
  n.id[s] = o.id[f];
  const float4 p = pos32_old[f];
  pos32[s] = p;
  pos32_next[s] = p;
  const double4 d = posd_old[f];
  posd[s] = d;
  posd_next[s] = d;
}

// A wavefront per new cluster of a kept box: its box, centre, tower and cluster-pair row from the old
// cluster, the j-cluster indices of the row shifted with the region; an added box's clusters get empty
// rows until its build.
constexpr int kMoveThreads = 256;

__global__ __launch_bounds__(kMoveThreads) void k_move_clusters(
    int64_t nclusters, const int32_t *__restrict__ cluster_box, const int64_t *__restrict__ cbase,
    const int64_t *__restrict__ src_cbase, const int64_t *__restrict__ tower_delta, const double *__restrict__ bbox_old,
    const double4 *__restrict__ c0_old, const int32_t *__restrict__ count_old, const int32_t *__restrict__ ctower_old,
    const int4 *__restrict__ list_old, int32_t cap, double *__restrict__ bbox, double4 *__restrict__ c0,
    int32_t *__restrict__ count, int32_t *__restrict__ ctower, int4 *__restrict__ list) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t c = (int64_t)blockIdx.x * (kMoveThreads / kWave) + threadIdx.x / kWave;
  if (c >= nclusters) return;
  const int32_t k = cluster_box[c];
  const int64_t sc = src_cbase[k];
  if (sc < 0) {
    if (lane == 0) count[c] = 0;
    return;
  }
  const int64_t f = c + (sc - cbase[k]);
  const int32_t dc = (int32_t)(cbase[k] - sc);
  const int32_t n = count_old[f] < cap ? count_old[f] : cap;
  if (lane < 6) bbox[6 * c + lane] = bbox_old[6 * f + lane];
  if (lane == 0) {
    c0[c] = c0_old[f];
    count[c] = n;
    const int32_t t = ctower_old[f];
    ctower[c] = t >= 0 ? (int32_t)(t + tower_delta[k]) : -1;
  }
  for (int32_t e = lane; e < n; e += kWave) {
    int4 v = list_old[f * (int64_t)cap + e];
    v.x = (v.x & ~kClusterMask) | ((v.x & kClusterMask) + dc);
    list[c * (int64_t)cap + e] = v;
  }
}

// The same move for the member rows of the FP32 cluster kernel ({j slot, entry}: the slot shifted with the
// region, the entry relative to the cluster's own row); an added box's clusters get empty rows until its
// build.
__global__ __launch_bounds__(kMoveThreads) void k_move_xrows(int64_t nclusters, const int32_t *__restrict__ cluster_box,
                                                             const int64_t *__restrict__ cbase,
                                                             const int64_t *__restrict__ src_cbase, int32_t M,
                                                             const int32_t *__restrict__ count_old,
                                                             const int2 *__restrict__ xrow_old, int32_t stride,
                                                             int32_t *__restrict__ count, int2 *__restrict__ xrow) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t c = (int64_t)blockIdx.x * (kMoveThreads / kWave) + threadIdx.x / kWave;
  if (c >= nclusters) return;
  const int32_t k = cluster_box[c];
  const int64_t sc = src_cbase[k];
  if (sc < 0) {
    if (lane == 0) count[c] = 0;
    return;
  }
  const int64_t f = c + (sc - cbase[k]);
  const int32_t ds = (int32_t)(cbase[k] - sc) * M;
  const int32_t n = count_old[f] < stride ? count_old[f] : stride;
  if (lane == 0) count[c] = n;
  for (int32_t e = lane; e < n; e += kWave) {
    int2 v = xrow_old[f * (int64_t)stride + e];
    v.x += ds;
    xrow[c * (int64_t)stride + e] = v;
  }
}

// Added particle k (upload index first + k) into its box's region: box b by its upload index
// (pbase[b0 ..] searched), slot cbase[b] M + (index - pbase[b]).
__global__ void k_add_to_regions(int64_t first, int64_t count, int32_t b0, int32_t nboxes, int32_t M,
                                 const int64_t *__restrict__ pbase, const int64_t *__restrict__ cbase, StateView n,
                                 const double *__restrict__ in, int32_t have_v, const int32_t *__restrict__ type,
                                 const int64_t *__restrict__ id) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int64_t i = first + k;
  int32_t lo = b0, hi = nboxes - 1;  // last box with pbase[b] <= i
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (pbase[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int64_t s = cbase[lo] * M + (i - pbase[lo]);
  const double x = in[k], y = in[count + k], z = in[2 * count + k];
  n.x[s] = x; n.y[s] = y; n.z[s] = z;
  n.rx[s] = x; n.ry[s] = y; n.rz[s] = z;
  n.vx[s] = have_v ? in[3 * count + k] : 0.;
  n.vy[s] = have_v ? in[4 * count + k] : 0.;
  n.vz[s] = have_v ? in[5 * count + k] : 0.;
  n.fx[s] = 0.; n.fy[s] = 0.; n.fz[s] = 0.;
  n.ofx[s] = 0.; n.ofy[s] = 0.; n.ofz[s] = 0.;
  n.type[s] = type ? type[k] : 0;
  n.orig[s] = (int32_t)i;
  n.id[s] = id ? id[k] : i;
}

void ensemble_edit_regions(Ctx &c, const std::vector<int32_t> &src_box, const std::vector<int64_t> &old_pstart,
                           int64_t kept, int64_t added, const double *x, const double *y, const double *z,
                           const double *vx, const double *vy, const double *vz, const int32_t *type,
                           const int64_t *id) {
  // the old layout's tables, before ensemble_geometry replaces them
  const std::vector<int64_t> old_cbase = c.box_cbase_h, old_tbase = c.box_tbase_h;
  const int32_t M = c.M, nb = c.nboxes, nkeep = (int32_t)src_box.size();
  c.box_geom_M = 0;
  ensemble_geometry(c);
  const std::vector<int64_t> &cbase = c.box_cbase_h, &tbase = c.box_tbase_h;
  std::vector<int64_t> tab(3 * (size_t)nb, -1);  // src_cbase, tower_delta, orig_delta per new box
  std::vector<int64_t> pbase((size_t)nb + 1, 0);
  for (int32_t b = 0; b < nb; ++b) pbase[b + 1] = pbase[b] + c.box_count[b];
  for (int32_t k = 0; k < nkeep; ++k) {
    const int32_t b = src_box[k];
    // a kept box has the grid and region it had: same bounds, count and cluster size
    if (cbase[k + 1] - cbase[k] != old_cbase[b + 1] - old_cbase[b] || tbase[k + 1] - tbase[k] != old_tbase[b + 1] - old_tbase[b])
      throw HipError{hipErrorUnknown, "a kept box changed its region", __LINE__};
    tab[k] = old_cbase[b];
    tab[nb + k] = tbase[k] - old_tbase[b];
    tab[2 * (size_t)nb + k] = pbase[k] - old_pstart[b];
  }
  for (int32_t k = nkeep; k < nb; ++k) tab[nb + k] = tab[2 * (size_t)nb + k] = 0;
  c.scan_tmp.ensure(tab.size() + pbase.size());
  int64_t *d_tab = c.scan_tmp.p, *d_pbase = c.scan_tmp.p + tab.size();
  LJG_HIP(hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
  LJG_HIP(hipMemcpyAsync(d_pbase, pbase.data(), pbase.size() * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
  const int64_t ncl = c.box_clusters, ns = ncl * M;
  const int other = 1 - c.cur;
  c.st[other].ensure((size_t)ns);
  // the moved layout goes to the second buffers, which are then swapped in (no allocation once they fit)
  DevBuf<float4> &p32 = c.pos32_alt, &p32n = c.pos32n_alt;
  DevBuf<double4> &pd = c.posd_alt, &pdn = c.posdn_alt;
  DevBuf<double> &bb = c.bbox_alt;
  DevBuf<double4> &cc = c.c0_alt;
  DevBuf<int32_t> &cnt = c.lcount_alt, &ctw = c.ctower_alt;
  DevBuf<int4> &rows = c.list_tmp;
  const int32_t cap = c.list_row;
  p32.ensure((size_t)ns);
  p32n.ensure((size_t)ns);
  pd.ensure((size_t)ns);
  pdn.ensure((size_t)ns);
  bb.ensure(6 * (size_t)ncl);
  cc.ensure((size_t)ncl);
  cnt.ensure((size_t)ncl * kJclSplit + 1);  // the size every build asks for, so a partial one keeps it
  ctw.ensure((size_t)ncl);
  rows.ensure((size_t)ncl * cap + 1);
  hipLaunchKernelGGL(k_move_slots, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, M, c.cluster_box.p,
                     c.box_cbase.p, d_tab, d_tab + 2 * nb, view(c.st[c.cur]), view(c.st[other]), c.pos32.p, p32.p,
                     p32n.p, c.posd.p, pd.p, pdn.p);
  LJG_HIP(hipGetLastError());
  constexpr int per_block = kMoveThreads / kWave;
  hipLaunchKernelGGL(k_move_clusters, dim3(blocks_for(ncl, per_block)), dim3(kMoveThreads), 0, c.stream, ncl,
                     c.cluster_box.p, c.box_cbase.p, d_tab, d_tab + nb, c.bbox.p, c.c0.p, c.list_count.p,
                     c.cluster_tower.p, c.list.p, cap, bb.p, cc.p, cnt.p, ctw.p, rows.p);
  LJG_HIP(hipGetLastError());
  const bool xrows = c.xrows_valid && c.xstride == cap * M;
  if (xrows) {
    c.xrow_tmp.ensure((size_t)ncl * c.xstride + 1);
    c.xcount_tmp.ensure((size_t)ncl + 1);
    hipLaunchKernelGGL(k_move_xrows, dim3(blocks_for(ncl, per_block)), dim3(kMoveThreads), 0, c.stream, ncl,
                       c.cluster_box.p, c.box_cbase.p, d_tab, M, c.xcount.p, c.xrow.p, c.xstride, c.xcount_tmp.p,
                       c.xrow_tmp.p);
    LJG_HIP(hipGetLastError());
  }
  if (added > 0) {
    const bool have_v = vx && vy && vz;
    const size_t n = (size_t)added;
    c.part_v.ensure(6 * n);
    const double *src[6] = {x, y, z, vx, vy, vz};
    for (int k = 0; k < (have_v ? 6 : 3); ++k)
      LJG_HIP(hipMemcpyAsync(c.part_v.p + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    int32_t *d_type = nullptr;
    int64_t *d_id = nullptr;
    if (type) {
      c.rank.ensure(n);
      d_type = c.rank.p;
      LJG_HIP(hipMemcpyAsync(d_type, type, n * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    }
    if (id) {
      c.tower_items.ensure(2 * n);  // scratch: the sort input is rebuilt by the next build
      d_id = reinterpret_cast<int64_t *>(c.tower_items.p);
      LJG_HIP(hipMemcpyAsync(d_id, id, n * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
    }
    hipLaunchKernelGGL(k_add_to_regions, dim3(blocks_for(added, 256)), dim3(256), 0, c.stream, kept, added, nkeep,
                       nb, M, d_pbase, c.box_cbase.p, view(c.st[other]), c.part_v.p, (int32_t)have_v, d_type, d_id);
    LJG_HIP(hipGetLastError());
  }
  LJG_HIP(hipStreamSynchronize(c.stream));  // the host tables go out of scope
  std::swap(c.pos32, p32);
  std::swap(c.pos32_next, p32n);
  std::swap(c.posd, pd);
  std::swap(c.posd_next, pdn);
  std::swap(c.bbox, bb);
  std::swap(c.c0, cc);
  std::swap(c.list_count, cnt);
  std::swap(c.cluster_tower, ctw);
  std::swap(c.list, rows);
  if (xrows) {
    std::swap(c.xrow, c.xrow_tmp);
    std::swap(c.xcount, c.xcount_tmp);
  }
  c.xrows_valid = xrows;  // without them (another stride) the next build of the boxes makes them all again
  c.cur = other;
  c.nslots = ns;
  c.nclusters = ncl;
  c.regions_valid = true;
}

__global__ void k_fill_owner(int64_t n, const int64_t *__restrict__ start, int32_t nb, int32_t *__restrict__ owner) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t lo = 0, hi = nb - 1;  // last b with start[b] <= i
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (start[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  owner[i] = lo;
}

void fill_owner(Ctx &c, int32_t *owner, int64_t n, const int64_t *start, int32_t nb) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_owner, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, start, nb, owner);
  LJG_HIP(hipGetLastError());
}

void ensemble_geometry(Ctx &c) {
  if (c.box_geom_M == c.M) return;
  const double ilen = c.prm.cutoff + c.prm.skin;
  const int32_t nb = c.nboxes;
  std::vector<TowerGeom> tg((size_t)nb);
  std::vector<int64_t> tbase((size_t)nb + 1);
  tbase[0] = 0;
  for (int32_t b = 0; b < nb; ++b) {
    tg[b] = make_tower_geom(&c.box_lo[3 * b], &c.box_hi[3 * b], c.box_count[b], c.M, ilen);
    tbase[b + 1] = tbase[b] + (int64_t)tg[b].nt[0] * tg[b].nt[1];
  }
  const int64_t ntow = tbase[nb];
  if (ntow > INT32_MAX / 2) throw HipError{hipErrorInvalidValue, "too many towers in the ensemble", -1};
  // regions: ceil(count/M) clusters plus one per tower bound sum_t ceil(count_t/M) for any distribution
  std::vector<int64_t> cbase((size_t)nb + 1), pbase((size_t)nb + 1);
  cbase[0] = pbase[0] = 0;
  for (int32_t b = 0; b < nb; ++b) {
    cbase[b + 1] = cbase[b] + (c.box_count[b] + c.M - 1) / c.M + (tbase[b + 1] - tbase[b]);
    pbase[b + 1] = pbase[b] + c.box_count[b];
  }
  const int64_t ncl = cbase[nb];
  if (ncl > (int64_t)kClusterMask) throw HipError{hipErrorInvalidValue, "too many clusters in the ensemble", -1};
  c.box_tg.ensure((size_t)nb);
  c.box_tbase.ensure((size_t)nb + 1);
  c.box_cbase.ensure((size_t)nb + 1);
  c.box_pbase.ensure((size_t)nb + 1);
  c.tower_box.ensure((size_t)ntow);
  c.cluster_box.ensure((size_t)ncl);
  LJG_HIP(hipMemcpyAsync(c.box_tg.p, tg.data(), tg.size() * sizeof(TowerGeom), hipMemcpyHostToDevice, c.stream));
  LJG_HIP(hipMemcpyAsync(c.box_tbase.p, tbase.data(), tbase.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                         c.stream));
  LJG_HIP(hipMemcpyAsync(c.box_cbase.p, cbase.data(), cbase.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                         c.stream));
  LJG_HIP(hipMemcpyAsync(c.box_pbase.p, pbase.data(), pbase.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                         c.stream));
  fill_owner(c, c.tower_box.p, ntow, c.box_tbase.p, nb);
  fill_owner(c, c.cluster_box.p, ncl, c.box_cbase.p, nb);
  LJG_HIP(hipStreamSynchronize(c.stream));  // the host tables go out of scope
  c.box_ntowers = ntow;
  c.box_clusters = ncl;
  c.box_cbase_h = std::move(cbase);
  c.box_tbase_h = std::move(tbase);
  c.regions_valid = false;
  c.tg = tg[0];
  c.box_geom_M = c.M;
}

void ensemble_tower_hist(Ctx &c, const int32_t *active) {
  State &s = c.st[c.cur];
  const int64_t ni = active ? c.seg_nslots : c.nslots;
  const Segs sg = active ? c.seg_slots : Segs{nullptr, nullptr, 0};
  if (ni == 0) return;
  hipLaunchKernelGGL(k_tower_hist_boxes, dim3(blocks_for(ni, 256)), dim3(256), 0, c.stream, ni, sg, s.x.p, s.y.p,
                     s.z.p, s.orig.p, c.box_of.p, c.box_tg.p, c.box_wrap.p, c.box_tbase.p, active, c.tower_of.p,
                     c.rank.p, c.tower_count.p);
  LJG_HIP(hipGetLastError());
}

void ensemble_layout(Ctx &c, const int32_t *active) {
  constexpr int per_block = kBoxLayoutThreads / kWave;
  const int32_t nb = active ? c.seg_nboxes : c.nboxes;
  if (nb == 0) return;
  hipLaunchKernelGGL(k_box_layout, dim3(blocks_for(nb, per_block)), dim3(kBoxLayoutThreads), 0, c.stream, nb,
                     active ? c.seg_boxes : (const int64_t *)nullptr, c.box_tbase.p, c.box_pbase.p, c.box_cbase.p,
                     active, c.tower_count.p, c.M, c.tower_start.p, c.ctow_count.p, c.ctow_start.p, c.tstats.p);
  LJG_HIP(hipGetLastError());
}

void ensemble_region_prep(Ctx &c, const int32_t *active) {
  const int64_t ni = active ? c.seg_nslots : c.box_clusters * c.M;
  const Segs sg = active ? c.seg_slots : Segs{nullptr, nullptr, 0};
  if (ni == 0) return;
  hipLaunchKernelGGL(k_region_prep, dim3(blocks_for(ni, 256)), dim3(256), 0, c.stream, ni, sg, c.M, c.cluster_box.p,
                     active, c.perm.p, c.cluster_tower.p);
  LJG_HIP(hipGetLastError());
}

void ensemble_restride_lists(Ctx &c, const int32_t *active, int32_t oldcap, int32_t newcap) {
  const int64_t ncl = c.box_clusters;
  c.list_tmp.ensure((size_t)ncl * newcap + 1);
  constexpr int per_block = kRestrideThreads / kWave;
  hipLaunchKernelGGL(k_restride_rows, dim3(blocks_for(ncl, per_block)), dim3(kRestrideThreads), 0, c.stream, ncl,
                     c.cluster_box.p, active, c.list_count.p, c.list.p, oldcap, c.list_tmp.p, newcap);
  LJG_HIP(hipGetLastError());
  std::swap(c.list, c.list_tmp);
}

void ensemble_wrap(Ctx &c) {
  hipLaunchKernelGGL(k_wrap_boxes, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots,
                     view(c.st[c.cur]), c.box_of.p, c.box_wrap.p);
  LJG_HIP(hipGetLastError());
}

void ensemble_box_globals(Ctx &c) {
  c.box_glob.ensure(3 * (size_t)c.nboxes);
  constexpr int per_block = kBoxReduceThreads / kWave;
  hipLaunchKernelGGL(k_box_reduce, dim3(blocks_for(c.nboxes, per_block)), dim3(kBoxReduceThreads), 0, c.stream,
                     c.nboxes, c.box_cbase.p, c.cl_glob.p, c.box_glob.p);
  LJG_HIP(hipGetLastError());
}

}  // namespace ljg
