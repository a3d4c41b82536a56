// ljg_vcl.hip -- Verlet-cluster build on the device (periodic box, shift vectors instead of halo copies):
//   tower of every particle (ClusterTowerBlock2D.h:220-245), counting sort into towers, per-tower z sort
//   in LDS (ClusterTower.h:93; ties by particle id), clusters of M consecutive particles padded with
//   dummies (ClusterTower.h:95-103), cluster bounding boxes (Cluster.h:135-149) and the cluster-pair
//   lists (VerletClusterListsRebuilder.h:237-409, boxDistanceSquared ArrayMath.h:697-707).
// Compiled with -ffp-contract=off so every FP64 comparison rounds exactly like the reference expression.
#include <cmath>

#include "ljg_internal.h"

namespace ljg {

TowerGeom make_tower_geom(const double lo[3], const double hi[3], int64_t n, int32_t M, double ilen) {
  TowerGeom g{};
  double L[3];
  for (int d = 0; d < 3; ++d) {
    L[d] = hi[d] - lo[d];
    g.box_min[d] = lo[d];
    g.L[d] = L[d];
  }
  const double density = (double)n / (L[0] * L[1] * L[2]);
  const double side0 = std::cbrt((double)M / density);
  for (int d = 0; d < 2; ++d) {
    long t = (long)std::ceil(L[d] / side0);
    if (t < 1) t = 1;
    g.nt[d] = (int32_t)t;
    g.side[d] = L[d] / (double)t;
    g.recip[d] = 1. / g.side[d];
  }
  const long r0 = (long)std::ceil(ilen / g.side[0]), r1 = (long)std::ceil(ilen / g.side[1]);
  g.R = (int32_t)(r0 > r1 ? r0 : r1);
  g.ilen = ilen;
  return g;
}

// Also wraps the positions into the box first (the rebuild's wrap, RegularGridDecomposition.cpp:563-587).
__global__ void k_tower_hist(int64_t nslots, double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                             const int32_t *__restrict__ orig, TowerGeom g, WrapArgs w,
                             int32_t *__restrict__ tower_of, int32_t *__restrict__ rank, int32_t *__restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < nslots && orig[i] >= 0;
  double px = 0., py = 0.;
  if (active) {
    const double ox = x[i], oy = y[i], oz = z[i];
    px = wrap_coord(ox, 0, w);
    py = wrap_coord(oy, 1, w);
    const double pz = wrap_coord(oz, 2, w);
    if (px != ox) x[i] = px;
    if (py != oy) y[i] = py;
    if (pz != oz) z[i] = pz;
  }
  const int32_t t = active ? tower_index(g, px, py) : -1;
  const int32_t r = wave_aggregated_rank(t, active, count);
  if (i < nslots) {
    tower_of[i] = t;
    rank[i] = r;
  }
}

// Also copies each particle's z and id next to its slot, so that the register sorts read them contiguously.
__global__ void k_tower_scatter(int64_t nitems, Segs sg, const int32_t *__restrict__ tower_of,
                                const int32_t *__restrict__ rank, const int64_t *__restrict__ start,
                                const double *__restrict__ z, const int64_t *__restrict__ id,
                                int32_t *__restrict__ items, double *__restrict__ items_z,
                                int64_t *__restrict__ items_id) {
  const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nitems) return;
  const int64_t i = seg_index(sg, it);
  const int32_t t = tower_of[i];
  if (t < 0) return;
  const int64_t k = start[t] + rank[i];
  items[k] = (int32_t)i;
  items_z[k] = z[i];
  items_id[k] = id[i];
}

// One workgroup: exclusive scans of the tower sizes (particle offsets) and of the clusters per tower,
// ceil(count / M), with the totals at index ntowers; stats = {particles, longest tower, clusters, shortest
// non-empty tower}.  Each thread scans a contiguous run of towers, then the run totals are scanned in LDS.
constexpr int kLayoutThreads = 1024;

__global__ __launch_bounds__(kLayoutThreads) void k_tower_layout(int64_t ntowers, const int32_t *__restrict__ count,
                                                                 int32_t M, int64_t *__restrict__ start,
                                                                 int32_t *__restrict__ ccount,
                                                                 int64_t *__restrict__ cstart,
                                                                 unsigned long long *__restrict__ stats) {
  __shared__ int64_t sp[kLayoutThreads], sc[kLayoutThreads];
  __shared__ int32_t sm[kLayoutThreads], sn[kLayoutThreads];
  const int tid = threadIdx.x;
  const int64_t per = (ntowers + kLayoutThreads - 1) / kLayoutThreads;
  const int64_t b = tid * per, e = b + per < ntowers ? b + per : ntowers;
  int64_t np = 0, nc = 0;
  int32_t mx = 0, mn = INT32_MAX;
  for (int64_t t = b; t < e; ++t) {
    const int32_t k = count[t];
    np += k;
    nc += (k + M - 1) / M;
    mx = k > mx ? k : mx;
    if (k > 0) mn = k < mn ? k : mn;
  }
  sp[tid] = np;
  sc[tid] = nc;
  sm[tid] = mx;
  sn[tid] = mn;
  __syncthreads();
  // inclusive scan of the run totals (Hillis-Steele), maximum by tree
  for (int off = 1; off < kLayoutThreads; off <<= 1) {
    const int64_t ap = tid >= off ? sp[tid - off] : 0, ac = tid >= off ? sc[tid - off] : 0;
    __syncthreads();
    sp[tid] += ap;
    sc[tid] += ac;
    __syncthreads();
  }
  for (int s = kLayoutThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      sm[tid] = sm[tid + s] > sm[tid] ? sm[tid + s] : sm[tid];
      sn[tid] = sn[tid + s] < sn[tid] ? sn[tid + s] : sn[tid];
    }
    __syncthreads();
  }
  int64_t op = sp[tid] - np, oc = sc[tid] - nc;
  for (int64_t t = b; t < e; ++t) {
    const int32_t k = count[t];
    const int32_t kc = (k + M - 1) / M;
    start[t] = op;
    cstart[t] = oc;
    ccount[t] = kc;
    op += k;
    oc += kc;
  }
  if (tid == kLayoutThreads - 1) {
    start[ntowers] = sp[tid];
    cstart[ntowers] = sc[tid];
    stats[0] = (unsigned long long)sp[tid];
    stats[2] = (unsigned long long)sc[tid];
  }
  if (tid == 0) {
    stats[1] = (unsigned long long)sm[0];
    stats[3] = (unsigned long long)(sn[0] == INT32_MAX ? 0 : sn[0]);
  }
}

// The same layout for grids of many towers (ensembles of boxes), in three passes so that every thread
// handles kLayoutPer towers: per-block totals, a scan of the block totals (one workgroup), and the
// per-tower offsets.  The totals and statistics are those of k_tower_layout.
constexpr int kLayoutBlock = 256;
constexpr int kLayoutPer = 16;
constexpr int64_t kLayoutChunk = (int64_t)kLayoutBlock * kLayoutPer;
constexpr int64_t kLayoutOneBlockMax = 16384;  // towers up to which the single-workgroup kernel is used

__global__ __launch_bounds__(kLayoutBlock) void k_layout_sums(int64_t ntowers, const int32_t *__restrict__ count,
                                                              int32_t M, int64_t *__restrict__ bsum,
                                                              int32_t *__restrict__ bmax) {
  __shared__ int64_t sp[kLayoutBlock], sc[kLayoutBlock];
  __shared__ int32_t sm[kLayoutBlock];
  const int tid = threadIdx.x;
  const int64_t b = (int64_t)blockIdx.x * kLayoutChunk + (int64_t)tid * kLayoutPer;
  const int64_t e = b + kLayoutPer < ntowers ? b + kLayoutPer : ntowers;
  int64_t np = 0, nc = 0;
  int32_t mx = 0;
  for (int64_t t = b; t < e; ++t) {
    const int32_t k = count[t];
    np += k;
    nc += (k + M - 1) / M;
    mx = k > mx ? k : mx;
  }
  sp[tid] = np;
  sc[tid] = nc;
  sm[tid] = mx;
  __syncthreads();
  for (int s = kLayoutBlock / 2; s > 0; s >>= 1) {
    if (tid < s) {
      sp[tid] += sp[tid + s];
      sc[tid] += sc[tid + s];
      sm[tid] = sm[tid + s] > sm[tid] ? sm[tid + s] : sm[tid];
    }
    __syncthreads();
  }
  if (tid == 0) {
    bsum[2 * blockIdx.x] = sp[0];
    bsum[2 * blockIdx.x + 1] = sc[0];
    bmax[blockIdx.x] = sm[0];
  }
}

// Exclusive scan of the block totals in place; totals at index ntowers of start/cstart and the stats.
__global__ __launch_bounds__(kLayoutThreads) void k_layout_top(int64_t nblocks, int64_t ntowers,
                                                               int64_t *__restrict__ bsum,
                                                               const int32_t *__restrict__ bmax,
                                                               int64_t *__restrict__ start,
                                                               int64_t *__restrict__ cstart,
                                                               unsigned long long *__restrict__ stats) {
  __shared__ int64_t sp[kLayoutThreads], sc[kLayoutThreads];
  __shared__ int32_t sm[kLayoutThreads];
  const int tid = threadIdx.x;
  const int64_t per = (nblocks + kLayoutThreads - 1) / kLayoutThreads;
  const int64_t b = tid * per, e = b + per < nblocks ? b + per : nblocks;
  int64_t np = 0, nc = 0;
  int32_t mx = 0;
  for (int64_t k = b; k < e; ++k) {
    np += bsum[2 * k];
    nc += bsum[2 * k + 1];
    mx = bmax[k] > mx ? bmax[k] : mx;
  }
  sp[tid] = np;
  sc[tid] = nc;
  sm[tid] = mx;
  __syncthreads();
  for (int off = 1; off < kLayoutThreads; off <<= 1) {
    const int64_t ap = tid >= off ? sp[tid - off] : 0, ac = tid >= off ? sc[tid - off] : 0;
    __syncthreads();
    sp[tid] += ap;
    sc[tid] += ac;
    __syncthreads();
  }
  for (int s = kLayoutThreads / 2; s > 0; s >>= 1) {
    if (tid < s) sm[tid] = sm[tid + s] > sm[tid] ? sm[tid + s] : sm[tid];
    __syncthreads();
  }
  int64_t op = sp[tid] - np, oc = sc[tid] - nc;
  for (int64_t k = b; k < e; ++k) {
    const int64_t kp = bsum[2 * k], kc = bsum[2 * k + 1];
    bsum[2 * k] = op;
    bsum[2 * k + 1] = oc;
    op += kp;
    oc += kc;
  }
  if (tid == kLayoutThreads - 1) {
    start[ntowers] = sp[tid];
    cstart[ntowers] = sc[tid];
    stats[0] = (unsigned long long)sp[tid];
    stats[2] = (unsigned long long)sc[tid];
  }
  if (tid == 0) {
    stats[1] = (unsigned long long)sm[0];
    stats[3] = 0ull;  // the shortest tower is not tracked here
  }
}

__global__ __launch_bounds__(kLayoutBlock) void k_layout_apply(int64_t ntowers, const int32_t *__restrict__ count,
                                                               int32_t M, const int64_t *__restrict__ boff,
                                                               int64_t *__restrict__ start,
                                                               int32_t *__restrict__ ccount,
                                                               int64_t *__restrict__ cstart) {
  __shared__ int64_t sp[kLayoutBlock], sc[kLayoutBlock];
  const int tid = threadIdx.x;
  const int64_t b = (int64_t)blockIdx.x * kLayoutChunk + (int64_t)tid * kLayoutPer;
  const int64_t e = b + kLayoutPer < ntowers ? b + kLayoutPer : ntowers;
  int64_t np = 0, nc = 0;
  for (int64_t t = b; t < e; ++t) {
    const int32_t k = count[t];
    np += k;
    nc += (k + M - 1) / M;
  }
  sp[tid] = np;
  sc[tid] = nc;
  __syncthreads();
  for (int off = 1; off < kLayoutBlock; off <<= 1) {
    const int64_t ap = tid >= off ? sp[tid - off] : 0, ac = tid >= off ? sc[tid - off] : 0;
    __syncthreads();
    sp[tid] += ap;
    sc[tid] += ac;
    __syncthreads();
  }
  int64_t op = boff[2 * blockIdx.x] + sp[tid] - np, oc = boff[2 * blockIdx.x + 1] + sc[tid] - nc;
  for (int64_t t = b; t < e; ++t) {
    const int32_t k = count[t];
    const int32_t kc = (k + M - 1) / M;
    start[t] = op;
    cstart[t] = oc;
    ccount[t] = kc;
    op += k;
    oc += kc;
  }
}

// One workgroup per tower: bitonic sort of (z, id) in LDS, then the tower's slots of the cluster layout
// are written: src[new slot] = old slot, -1 for padding; cluster_tower for the tower's clusters.  A tower
// holds count[t] particles, its items from start[t], its clusters [cstart[t], cstart[t] + ccount[t]).
// Two instances: CAP = kSortSmall takes the towers up to that size (12 KB of LDS, so many workgroups
// share a CU), CAP = kTowerSortCap the rare larger ones; each skips the towers of the other.
constexpr int kSortThreads = 256;
constexpr int kSortSmall = 1024;

__device__ inline bool zid_less(double za, int32_t sa, double zb, int32_t sb, const int64_t *__restrict__ id) {
  if (sb < 0) return sa >= 0;  // padding sorts last
  if (sa < 0) return false;
  if (za < zb) return true;
  if (za > zb) return false;
  return id[sa] < id[sb];
}

// Towers of up to min_n particles are left to k_tower_sort_wave (min_n = 0 when it does not run).
template <int CAP>
__global__ __launch_bounds__(kSortThreads) void k_tower_sort(
    int64_t ntowers, Segs sg, const int64_t *__restrict__ start, const int32_t *__restrict__ count,
    const int32_t *__restrict__ items, const double *__restrict__ z, const int64_t *__restrict__ id,
    const int64_t *__restrict__ cstart, const int32_t *__restrict__ ccount, int32_t M, int32_t *__restrict__ src,
    int32_t *__restrict__ cluster_tower, unsigned long long *__restrict__ err, int32_t min_n) {
  __shared__ double sz[CAP];
  __shared__ int32_t ss[CAP];
  if ((int64_t)blockIdx.x >= ntowers) return;
  const int64_t t = seg_index(sg, blockIdx.x);
  const int64_t b = start[t];
  const int n = count[t];
  if (n <= min_n) return;                           // sorted in registers by k_tower_sort_wave
  if (CAP > kSortSmall && n <= kSortSmall) return;  // sorted by the small instance
  if (CAP == kSortSmall && n > kSortSmall) return;  // left to the large instance
  if (n > CAP) {
    if (threadIdx.x == 0) err[0] = 1ull;
    return;
  }
  int P = 1;
  while (P < n) P <<= 1;
  for (int k = threadIdx.x; k < P; k += kSortThreads) {
    if (k < n) {
      const int32_t s = items[b + k];
      ss[k] = s;
      sz[k] = z[s];
    } else {
      ss[k] = -1;
      sz[k] = 0.;
    }
  }
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = threadIdx.x; k < P / 2; k += kSortThreads) {
        const int lo = 2 * k - (k & (stride - 1));
        const int hi = lo + stride;
        const bool up = ((lo & size) == 0);
        const double za = sz[lo], zb = sz[hi];
        const int32_t sa = ss[lo], sb = ss[hi];
        const bool swap = up ? zid_less(zb, sb, za, sa, id) : zid_less(za, sa, zb, sb, id);
        if (swap) {
          sz[lo] = zb; sz[hi] = za;
          ss[lo] = sb; ss[hi] = sa;
        }
      }
      __syncthreads();
    }
  }
  const int64_t c0 = cstart[t], c1 = c0 + ccount[t];
  const int64_t nslot = (c1 - c0) * M;
  for (int64_t k = threadIdx.x; k < nslot; k += kSortThreads) src[c0 * M + k] = k < n ? ss[k] : -1;
  for (int64_t c = c0 + threadIdx.x; c < c1; c += kSortThreads) cluster_tower[c] = (int32_t)t;
}

// Short towers (small boxes): one wavefront per tower sorts (z, id) in registers with a bitonic network
// over 64 E elements, element e * 64 + lane in slot e of a lane (padding sorts last), and writes the same
// cluster layout as k_tower_sort.  Keys are distinct (ids are), so every sorter produces the one sorted
// order.  The E = 1 instance takes towers of up to 64 particles, E = 4 those of 65 to 256, E = 8 those
// of 257 to 512.
constexpr int kSortWaveThreads = 256;

// The register sort carries two words per element: its z rounded to FP32 and its index k in the tower's
// item list.  Rounding is monotone, so a smaller FP32 z means a smaller z; equal FP32 z (rare) are
// ordered by the exact (z, id) read back from the item list, in a branch a wavefront takes only when one
// of its lanes holds such a pair.  Indices k < 0 are padding and sort last.
struct TowerItems {
  const double *z;
  const int64_t *id;
};

__device__ inline void zk_less2(float za, int32_t ka, float zb, int32_t kb, const TowerItems &ti, bool &ab, bool &ba) {
  ab = (ka >= 0) & ((kb < 0) | (za < zb));
  ba = (kb >= 0) & ((ka < 0) | (zb < za));
  const bool tie = (ka >= 0) & (kb >= 0) & (za == zb);
  if (__builtin_amdgcn_ballot_w64(tie) != 0ull) {
    if (tie) {
      const double da = ti.z[ka], db = ti.z[kb];
      const int64_t ia = ti.id[ka], ib = ti.id[kb];
      ab = (da < db) | ((da == db) & (ia < ib));
      ba = !ab;
    }
  }
}

template <int E, int SIZE, int STRIDE>
__device__ inline void bitonic2_stage(float (&zv)[E], int32_t (&kv)[E], int lane, const TowerItems &ti) {
  if constexpr (STRIDE >= kWave) {
    constexpr int ks = STRIDE / kWave;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if ((e & ks) != 0) continue;
      const int e2 = e + ks;
      const bool up = ((e * kWave + lane) & SIZE) == 0;
      bool lt, gt;  // zv[e2] before zv[e], and the reverse
      zk_less2(zv[e2], kv[e2], zv[e], kv[e], ti, lt, gt);
      const bool swap = up ? lt : gt;
      const float za = zv[e], zb = zv[e2];
      const int32_t ka = kv[e], kb = kv[e2];
      zv[e] = swap ? zb : za; zv[e2] = swap ? za : zb;
      kv[e] = swap ? kb : ka; kv[e2] = swap ? ka : kb;
    }
  } else {
    const bool lower = (lane & STRIDE) == 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float zo = __shfl_xor(zv[e], STRIDE, kWave);
      const int32_t ko = __shfl_xor(kv[e], STRIDE, kWave);
      const bool up = ((e * kWave + lane) & SIZE) == 0;
      bool o_first, me_first;
      zk_less2(zo, ko, zv[e], kv[e], ti, o_first, me_first);
      const bool take = (lower == up) ? o_first : me_first;
      zv[e] = take ? zo : zv[e];
      kv[e] = take ? ko : kv[e];
    }
  }
}

template <int E, int SIZE, int STRIDE>
__device__ inline void bitonic2_strides(float (&zv)[E], int32_t (&kv)[E], int lane, const TowerItems &ti) {
  bitonic2_stage<E, SIZE, STRIDE>(zv, kv, lane, ti);
  if constexpr (STRIDE > 1) bitonic2_strides<E, SIZE, STRIDE / 2>(zv, kv, lane, ti);
}

template <int E, int SIZE>
__device__ inline void bitonic2_sizes(float (&zv)[E], int32_t (&kv)[E], int lane, const TowerItems &ti) {
  bitonic2_strides<E, SIZE, SIZE / 2>(zv, kv, lane, ti);
  if constexpr (SIZE < kWave * E) bitonic2_sizes<E, SIZE * 2>(zv, kv, lane, ti);
}

template <int E>
__global__ __launch_bounds__(kSortWaveThreads) void k_tower_sort_wave(
    int64_t ntowers, Segs sg, const int64_t *__restrict__ start, const int32_t *__restrict__ count,
    const int32_t *__restrict__ items, const double *__restrict__ items_z, const int64_t *__restrict__ items_id,
    const int64_t *__restrict__ cstart, const int32_t *__restrict__ ccount, int32_t M, int32_t *__restrict__ src,
    int32_t *__restrict__ cluster_tower, unsigned long long *__restrict__ over) {
  constexpr int P = kWave * E;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w = (int64_t)blockIdx.x * (kSortWaveThreads / kWave) + threadIdx.x / kWave;
  if (w >= ntowers) return;
  const int64_t t = seg_index(sg, w);
  const int64_t b = start[t];
  const int n = count[t];
  constexpr int LO = E == 1 ? 0 : E == 4 ? kWave : 4 * kWave;  // the next shorter instance's reach
  static_assert(E == 1 || E == 4 || E == 8, "instances cover up to 64, 65-256 and 257-512 particles");
  // (over: the longest instance launched reports a tower beyond its reach, which no launched sort takes)
  if (n > P && over && lane == 0) over[0] = 1ull;
  if (n > P || n <= LO) return;  // another instance sorts this tower
  float zv[E];
  int32_t kv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = e * kWave + lane;
    kv[e] = k < n ? k : -1;
    zv[e] = k < n ? (float)items_z[b + k] : 0.f;
  }
  const TowerItems ti{items_z + b, items_id + b};
  bitonic2_sizes<E, 2>(zv, kv, lane, ti);
  const int64_t c0 = cstart[t], c1 = c0 + ccount[t];
  const int nslot = (int)((c1 - c0) * M);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = e * kWave + lane;
    if (k < nslot) src[c0 * M + k] = kv[e] >= 0 ? items[b + kv[e]] : -1;
  }
  if (P + lane < nslot) src[c0 * M + P + lane] = -1;
  for (int64_t c = c0 + lane; c < c1; c += kWave) cluster_tower[c] = (int32_t)t;
}

// Towers of 65 to 256 (E = 2) and 257 to 512 particles (E = 4): two wavefronts per tower, each sorting
// one half of up to 64 E in registers as k_tower_sort_wave<E> does, then a merge by rank: an element's
// place is its index in its half plus the number of the other half's elements before it (a binary
// search in the other half, held in LDS; equal FP32 z are ordered by the exact z and id).  Twice the
// wavefronts of a single-wave sort of the tower, each with half the network.
constexpr int kSortPairThreads = 2 * kWave;

template <int E>
__global__ __launch_bounds__(kSortPairThreads) void k_tower_sort_pair(
    int64_t ntowers, Segs sg, const int64_t *__restrict__ start, const int32_t *__restrict__ count,
    const int32_t *__restrict__ items, const double *__restrict__ items_z, const int64_t *__restrict__ items_id,
    const int64_t *__restrict__ cstart, const int32_t *__restrict__ ccount, int32_t M, int32_t *__restrict__ src,
    int32_t *__restrict__ cluster_tower, unsigned long long *__restrict__ over) {
  static_assert(E == 2 || E == 4, "pairs cover 65-256 and 257-512 particles");
  constexpr int P = kWave * E;             // each half up to 64 E
  constexpr int LO = E == 2 ? kWave : 4 * kWave;  // the next shorter instance's reach
  __shared__ float s_z[2][P];
  __shared__ int32_t s_k[2][P];
  const int lane = threadIdx.x & (kWave - 1), half = (int)threadIdx.x / kWave;
  if ((int64_t)blockIdx.x >= ntowers) return;  // the whole workgroup: one tower
  const int64_t t = seg_index(sg, blockIdx.x);
  const int64_t b = start[t];
  const int n = count[t];
  if (n > 2 * P && over && threadIdx.x == 0) over[0] = 1ull;
  if (n > 2 * P || n <= LO) return;  // another instance sorts this tower
  const int h = (n + 1) / 2;
  const int first = half ? h : 0, len = half ? n - h : h, olen = half ? h : n - h;
  float zf[E];
  int32_t kv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = e * kWave + lane;
    kv[e] = k < len ? first + k : -1;
    zf[e] = k < len ? (float)items_z[b + first + k] : 0.f;
  }
  const TowerItems ti{items_z + b, items_id + b};
  bitonic2_sizes<E, 2>(zf, kv, lane, ti);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    s_z[half][e * kWave + lane] = zf[e];
    s_k[half][e * kWave + lane] = kv[e];
  }
  __syncthreads();
  const float *oz = s_z[1 - half];
  const int32_t *ok = s_k[1 - half];
  const int64_t c0 = cstart[t], c1 = c0 + ccount[t];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = e * kWave + lane;
    if (idx >= len) continue;
    const float z = zf[e];
    const int32_t k = kv[e];
    int lo = 0, hi = olen;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const float zo = oz[mid];
      bool before = zo < z;
      if (zo == z) {
        const int32_t ko = ok[mid];
        const double da = ti.z[ko], db = ti.z[k];
        before = (da < db) | ((da == db) & (ti.id[ko] < ti.id[k]));
      }
      if (before) lo = mid + 1;
      else hi = mid;
    }
    src[c0 * M + idx + lo] = items[b + k];
  }
  const int nslot = (int)((c1 - c0) * M);
  for (int q = n + (int)threadIdx.x; q < nslot; q += kSortPairThreads) src[c0 * M + q] = -1;
  for (int64_t c = c0 + threadIdx.x; c < c1; c += kSortPairThreads) cluster_tower[c] = (int32_t)t;
}

// One lane per slot, the M lanes of a cluster reduce with shuffles (M divides the wavefront, and the
// cluster's slots are consecutive, so a cluster never straddles two wavefronts).  The members are sorted
// by z, so the first and last real members' z are the minimum and maximum.
// In an ensemble (cluster_box given) only the clusters of the boxes box_active marks (every box when it
// is null) are written, and an empty cluster of a region's tail gets an empty row (lcount = 0).
__global__ void k_cluster_bbox(int64_t nitems, Segs sg, int32_t M, const double *__restrict__ x,
                               const double *__restrict__ y, const double *__restrict__ z,
                               const int32_t *__restrict__ orig, double *__restrict__ bbox,
                               double4 *__restrict__ c0, const int32_t *__restrict__ cluster_box,
                               const int32_t *__restrict__ box_active, int32_t *__restrict__ lcount) {
  // segments are whole clusters, so the M lanes of a cluster stay together in a wavefront
  const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t s = it < nitems ? seg_index(sg, it) : -1;
  // a cluster's M slots are in one box: the test is uniform over the lanes that exchange values below
  const bool mine = s >= 0 && (box_active == nullptr || box_active[cluster_box[s / M]] != 0);
  const bool real = mine && orig[s] >= 0;
  const double inf = __builtin_inf();
  double xl = inf, xh = -inf, yl = inf, yh = -inf, zl = inf, zh = -inf;
  int nreal = real ? 1 : 0;
  if (real) {
    xl = xh = x[s];
    yl = yh = y[s];
    zl = zh = z[s];
  }
  for (int off = M >> 1; off > 0; off >>= 1) {
    xl = fmin(xl, __shfl_xor(xl, off, kWave));
    xh = fmax(xh, __shfl_xor(xh, off, kWave));
    yl = fmin(yl, __shfl_xor(yl, off, kWave));
    yh = fmax(yh, __shfl_xor(yh, off, kWave));
    zl = fmin(zl, __shfl_xor(zl, off, kWave));
    zh = fmax(zh, __shfl_xor(zh, off, kWave));
    nreal += __shfl_xor(nreal, off, kWave);
  }
  if (mine && s % M == 0) {
    const int64_t c = s / M;
    double *bb = bbox + 6 * c;
    bb[0] = xl; bb[1] = yl; bb[2] = zl; bb[3] = xh; bb[4] = yh; bb[5] = zh;
    c0[c] = make_double4((xl + xh) * 0.5, (yl + yh) * 0.5, (zl + zh) * 0.5, (double)nreal);
    if (lcount && nreal == 0) lcount[c] = 0;
  }
}

// Newton-3 lists keep only forward entries: a higher j-cluster, or the i-cluster itself at shift id 13 (its
// own member pairs) or above (one of each pair of opposite images); as the oracle's
// ljo_vcl_forward_only, VerletClusterListsRebuilder.h:339-354.
__device__ inline bool forward_entry(int64_t ci, int64_t cj, int sid) { return cj > ci || (cj == ci && sid >= 13); }

__device__ inline double box_dist2(const double *a, const double *b, double shx, double shy, double shz) {
  const double sh[3] = {shx, shy, shz};
  double s = 0.;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double bmin = b[d] + sh[d], bmax = b[3 + d] + sh[d];
    const double ab0 = a[d] - bmax;
    const double ba0 = bmin - a[3 + d];
    const double ab = ab0 > 0. ? ab0 : 0.;
    const double ba = ba0 > 0. ? ba0 : 0.;
    s += ab * ab;
    s += ba * ba;
  }
  return s;
}

// Cluster-pair lists, one workgroup (one wavefront) per i-tower: each neighbour tower's cluster boxes and
// reference points are staged in LDS once and shared by all i-clusters of the tower (one lane each).
// Rows have a fixed capacity (entries [ci*cap, ci*cap + count[ci])); the traversal order (oy, ox, sz, j)
// is the oracle's, so lists compare entry by entry.  The first j-cluster whose z-range can reach the
// i-cluster is a binary search on the monotone upper z-extents held in LDS.  count[ci] is the row's full
// length even when it exceeds cap, so the host can grow cap and rebuild if a row overflowed.
constexpr int kListThreads = 64;
constexpr int kListLds = 128;  // j-clusters staged per chunk

// JCL = true (the j-particle row build, first pass): entries are the j-cluster index only (int32), and a
// j-cluster that passes the box test under several images (small boxes, long clusters) is listed only at
// the first of them in traversal order, since the expansion pass tests minimum-image separations.
constexpr int kMaxM = 8;

__device__ inline int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Whether the j-cluster with box bj passed the box test at a visit before (oy, ox, sz) in the traversal
// order of k_cluster_lists_tower.  xy_repeats: the +-R tower window covers some tower twice.
__device__ inline bool passed_earlier(const double *bi, const double *bj, const TowerGeom &g, int tx, int ty,
                                      int oy, int ox, int sz, bool xy_repeats, double il2) {
  const double shx = floor_div(tx + ox, g.nt[0]) * g.L[0], shy = floor_div(ty + oy, g.nt[1]) * g.L[1];
  for (int s2 = -1; s2 < sz; ++s2)
    if (box_dist2(bi, bj, shx, shy, s2 * g.L[2]) <= il2) return true;
  if (!xy_repeats) return false;
  for (int oy2 = -g.R; oy2 <= oy; ++oy2) {
    if ((oy2 - oy) % g.nt[1] != 0) continue;
    const int aoy = oy2 < 0 ? -oy2 : oy2;
    const double dty = (double)(aoy - 1 > 0 ? aoy - 1 : 0) * g.side[1];
    for (int ox2 = -g.R; ox2 <= g.R; ++ox2) {
      if (oy2 == oy && ox2 >= ox) break;
      if ((ox2 - ox) % g.nt[0] != 0) continue;
      const int aox = ox2 < 0 ? -ox2 : ox2;
      const double dtx = (double)(aox - 1 > 0 ? aox - 1 : 0) * g.side[0];
      if (dtx * dtx + dty * dty > il2) continue;
      const double sx2 = floor_div(tx + ox2, g.nt[0]) * g.L[0], sy2 = floor_div(ty + oy2, g.nt[1]) * g.L[1];
      for (int s2 = -1; s2 <= 1; ++s2)
        if (box_dist2(bi, bj, sx2, sy2, s2 * g.L[2]) <= il2) return true;
    }
  }
  return false;
}

// LDS exchange among the lanes of one wavefront (the waves of a workgroup stage different towers).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// NS wavefronts per tower split the neighbour-tower visits (oy, ox) of the traversal into NS consecutive
// ranges; wave w writes its entries for i-cluster ci into sub-row w, [ci*cap + w*cap/NS, ...), with its
// length in count[ci*NS + w], so the sub-rows read in order give the single-wave traversal order.
// The cluster-entry lists (JCL = false, compared entry by entry with the oracle) use NS = 1.
// MB (an ensemble of boxes): tower t belongs to box tbox[t], whose grid is tgs[box] and whose towers
// start at tbase[box]; the traversal inside the box is the single-box one.
template <bool JCL, int NS, bool MB>
__global__ __launch_bounds__(kListThreads * NS, JCL ? 4 : 1) void k_cluster_lists_tower(
    int64_t ntowers, Segs sg, TowerGeom g1, const TowerGeom *__restrict__ tgs, const int64_t *__restrict__ tbase,
    const int32_t *__restrict__ tbox, const int32_t *__restrict__ box_active, const int64_t *__restrict__ cstart,
    const int32_t *__restrict__ ccount, const double *__restrict__ bbox, const double4 *__restrict__ c0,
    int32_t cap, int32_t *__restrict__ count, int4 *__restrict__ list, int32_t *__restrict__ jcl, int32_t n3) {
  // the j-cluster index pass (JCL) needs neither the reference points nor as deep a chunk, and the smaller
  // LDS footprint lets more workgroups share a CU
  constexpr int LDSN = JCL ? kListLds / 2 : kListLds;
  __shared__ double sbw[NS][LDSN * 6];
  __shared__ double4 scw[NS][JCL ? 1 : LDSN];
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  double *sb = sbw[wv];
  double4 *sc = scw[wv];
  if ((int64_t)blockIdx.x >= ntowers) return;
  const int64_t t = seg_index(sg, blockIdx.x);
  TowerGeom g = g1;
  int64_t tb = 0;
  if constexpr (MB) {
    const int32_t b = tbox[t];
    if (box_active && box_active[b] == 0) return;  // the whole workgroup: one tower
    g = tgs[b];
    tb = tbase[b];
  }
  const int64_t tl = t - tb;
  const int tx = (int)(tl % g.nt[0]), ty = (int)(tl / g.nt[0]);
  const int64_t ib = cstart[t], ie = ib + ccount[t];
  const double il2 = g.ilen * g.ilen;
  const double zwin = g.ilen * (1.0 + 1e-9) + 1e-9;
  const bool xy_repeats = g.nt[0] < 2 * g.R + 1 || g.nt[1] < 2 * g.R + 1;
  const int side = 2 * g.R + 1, nvis = side * side;
  const int vbeg = (int)((int64_t)nvis * wv / NS), vend = (int)((int64_t)nvis * (wv + 1) / NS);
  const int32_t subcap = cap / NS;
  for (int64_t base = ib; base < ie; base += kListThreads) {
    const int64_t ci = base + lane;
    const bool act = ci < ie;
    double bi[6] = {0., 0., 0., 0., 0., 0.};
    double4 ai = make_double4(0., 0., 0., 0.);
    if (act) {
#pragma unroll
      for (int k = 0; k < 6; ++k) bi[k] = bbox[6 * ci + k];
      ai = c0[ci];
    }
    const int64_t w0 = ci * (int64_t)cap + (int64_t)wv * subcap;
    int32_t cnt = 0;
    // The visits with work, in order: the neighbour tower of visit v within reach and not empty.
    auto next_visit = [&](int v0, int &vo, int &sxo, int &syo, int64_t &cbo, int64_t &ceo) {
      for (int v = v0; v < vend; ++v) {
        const int oy = v / side - g.R, ox = v % side - g.R;
        const int aoy = oy < 0 ? -oy : oy;
        const double dty = (double)(aoy - 1 > 0 ? aoy - 1 : 0) * g.side[1];
        const int aox = ox < 0 ? -ox : ox;
        const double dtx = (double)(aox - 1 > 0 ? aox - 1 : 0) * g.side[0];
        if (dtx * dtx + dty * dty > il2) continue;
        int jx = tx + ox, jy = ty + oy, sx = 0, sy = 0;
        while (jx < 0) { jx += g.nt[0]; --sx; }
        while (jx >= g.nt[0]) { jx -= g.nt[0]; ++sx; }
        while (jy < 0) { jy += g.nt[1]; --sy; }
        while (jy >= g.nt[1]) { jy -= g.nt[1]; ++sy; }
        const int64_t tj = tb + jx + (int64_t)jy * g.nt[0];
        const int64_t cb = cstart[tj], ce = cb + ccount[tj];
        if (cb == ce) continue;
        vo = v; sxo = sx; syo = sy; cbo = cb; ceo = ce;
        return true;
      }
      vo = vend;
      return false;
    };
    // JCL: the boxes of the next visit's tower (one chunk) are loaded into registers while this visit runs
    // from LDS, and stored to LDS when the next visit starts, so no visit waits for its tower's boxes
    constexpr int kPf = JCL ? 6 * LDSN / kWave : 1;
    double pf[kPf];
    auto prefetch = [&](int64_t pcb, int64_t pce) {
      const int pn = 6 * (int)(pce - pcb);
#pragma unroll
      for (int i = 0; i < kPf; ++i) {
        const int k = lane + i * kWave;
        pf[i] = k < pn ? bbox[6 * pcb + k] : 0.;
      }
    };
    int v = vbeg, sx = 0, sy = 0;
    int64_t cb = 0, ce = 0;
    bool have = next_visit(vbeg, v, sx, sy, cb, ce);
    if (JCL && have && ce - cb <= LDSN) prefetch(cb, ce);
    while (have) {
      int v2 = vend, sx2 = 0, sy2 = 0;
      int64_t cb2 = 0, ce2 = 0;
      const bool have2 = next_visit(v + 1, v2, sx2, sy2, cb2, ce2);
      const bool single = ce - cb <= LDSN;
      if (JCL && single) {
        const int pn = 6 * (int)(ce - cb);
        wave_sync();
#pragma unroll
        for (int i = 0; i < kPf; ++i) {
          const int k = lane + i * kWave;
          if (k < pn) sb[k] = pf[i];
        }
        wave_sync();
      }
      // (after this visit's boxes left the registers; a visit of several chunks stages its own)
      if (JCL && have2 && ce2 - cb2 <= LDSN) prefetch(cb2, ce2);
      const int oy = v / side - g.R, ox = v % side - g.R;
      for (int sz = -1; sz <= 1; ++sz) {
        const double shx = sx * g.L[0], shy = sy * g.L[1], shz = sz * g.L[2];
        const double zmin = bi[2] - zwin - shz;  // window in the j tower's own frame
        const double zmax = bi[5] + zwin - shz;
        for (int64_t jb = cb; jb < ce; jb += LDSN) {
          const int nj = (int)(ce - jb < LDSN ? ce - jb : LDSN);
          if (!single || (!JCL && sz == -1 && jb == cb)) {
            wave_sync();
            for (int k = lane; k < 6 * nj; k += kWave) sb[k] = bbox[6 * jb + k];
            if constexpr (!JCL)
              for (int k = lane; k < nj; k += kWave) sc[k] = c0[jb + k];
            wave_sync();
          }
          if (!act) continue;
          // every cluster's z-range lies in [box_min, box_min + L): a window outside it finds nothing
          if (zmax < g.box_min[2] || zmin > g.box_min[2] + g.L[2]) continue;
          // lower bound of zhi >= zmin in the chunk
          int lo = 0, hi = nj;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sb[6 * mid + 5] < zmin) lo = mid + 1;
            else hi = mid;
          }
          for (int k = lo; k < nj; ++k) {
            const double *bj = sb + 6 * k;
            if (bj[2] > zmax) break;
            if (box_dist2(bi, bj, shx, shy, shz) <= il2) {
              if constexpr (JCL) {
                if (!passed_earlier(bi, bj, g, tx, ty, oy, ox, sz, xy_repeats, il2)) {
                  if (cnt < subcap) jcl[w0 + cnt] = (int32_t)(jb + k);
                  ++cnt;
                }
              } else {
                const int sid = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
                if (n3 && !forward_entry(ci, jb + k, sid)) continue;
                if (cnt < subcap) {
                  const double4 b = sc[k];
                  const float dx = (float)((b.x + shx) - ai.x);
                  const float dy = (float)((b.y + shy) - ai.y);
                  const float dz = (float)((b.z + shz) - ai.z);
                  list[w0 + cnt] = make_int4((int32_t)(jb + k) | (sid << kShiftBits), __float_as_int(dx),
                                             __float_as_int(dy), __float_as_int(dz));
                }
                ++cnt;
              }
            }
          }
        }
      }
      v = v2; sx = sx2; sy = sy2; cb = cb2; ce = ce2; have = have2;
    }
    if (act) count[ci * NS + wv] = cnt;
  }
}

// One neighbour-tower visit v of an i-cluster's traversal (k_cluster_lists_tower's order: oy, ox, then sz,
// then the j-tower's clusters from the first whose z-range can reach the i-cluster), with the same tests:
// the number of entries it finds, which WRITE stores from row position pos on (below cap).
template <bool WRITE>
__device__ inline int32_t visit_entries(int v, int side, const TowerGeom &g, int64_t tb, int tx, int ty,
                                        const double *bi, const double4 &ai, const int64_t *__restrict__ cstart,
                                        const int32_t *__restrict__ ccount, const double *__restrict__ bbox,
                                        const double4 *__restrict__ c0, int32_t cap, int32_t pos,
                                        int4 *__restrict__ row, int64_t ci, int32_t n3) {
  const double il2 = g.ilen * g.ilen;
  const double zwin = g.ilen * (1.0 + 1e-9) + 1e-9;
  const int oy = v / side - g.R, ox = v % side - g.R;
  const int aoy = oy < 0 ? -oy : oy;
  const double dty = (double)(aoy - 1 > 0 ? aoy - 1 : 0) * g.side[1];
  const int aox = ox < 0 ? -ox : ox;
  const double dtx = (double)(aox - 1 > 0 ? aox - 1 : 0) * g.side[0];
  if (dtx * dtx + dty * dty > il2) return 0;
  int jx = tx + ox, jy = ty + oy, sx = 0, sy = 0;
  while (jx < 0) { jx += g.nt[0]; --sx; }
  while (jx >= g.nt[0]) { jx -= g.nt[0]; ++sx; }
  while (jy < 0) { jy += g.nt[1]; --sy; }
  while (jy >= g.nt[1]) { jy -= g.nt[1]; ++sy; }
  const int64_t tj = tb + jx + (int64_t)jy * g.nt[0];
  const int64_t cb = cstart[tj], ce = cb + ccount[tj];
  int32_t n = 0;
  for (int sz = -1; sz <= 1; ++sz) {
    const double shx = sx * g.L[0], shy = sy * g.L[1], shz = sz * g.L[2];
    const double zmin = bi[2] - zwin - shz;
    const double zmax = bi[5] + zwin - shz;
    // every cluster's z-range lies in [box_min, box_min + L): a window outside it finds nothing
    if (zmax < g.box_min[2] || zmin > g.box_min[2] + g.L[2]) continue;
    int64_t lo = cb, hi = ce;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (bbox[6 * mid + 5] < zmin) lo = mid + 1;
      else hi = mid;
    }
    for (int64_t k = lo; k < ce; ++k) {
      const double *bj = bbox + 6 * k;
      if (bj[2] > zmax) break;
      if (box_dist2(bi, bj, shx, shy, shz) <= il2) {
        const int sid = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
        if (n3 && !forward_entry(ci, k, sid)) continue;
        if (WRITE && pos + n < cap) {
          const double4 b = c0[k];
          const float dx = (float)((b.x + shx) - ai.x);
          const float dy = (float)((b.y + shy) - ai.y);
          const float dz = (float)((b.z + shz) - ai.z);
          row[pos + n] = make_int4((int32_t)k | (sid << kShiftBits), __float_as_int(dx), __float_as_int(dy),
                                   __float_as_int(dz));
        }
        ++n;
      }
    }
  }
  return n;
}

// Cluster-pair lists with one lane per i-cluster, for grids whose towers hold few clusters (small boxes,
// ensembles of them), where the wavefront per tower of k_cluster_lists_tower leaves most lanes idle.
// Each lane walks the traversal order of k_cluster_lists_tower visit by visit with the same tests, so the
// rows are identical entry by entry; the j-tower's boxes are read from memory (lanes of one tower read the
// same lines).
constexpr int kListLaneThreads = 256;

template <bool MB>
__global__ __launch_bounds__(kListLaneThreads) void k_cluster_lists_lane(
    int64_t nclusters, Segs sg, TowerGeom g1, const TowerGeom *__restrict__ tgs, const int64_t *__restrict__ tbase,
    const int32_t *__restrict__ tbox, const int32_t *__restrict__ cluster_box, const int32_t *__restrict__ box_active,
    const int32_t *__restrict__ cluster_tower, const int64_t *__restrict__ cstart, const int32_t *__restrict__ ccount,
    const double *__restrict__ bbox, const double4 *__restrict__ c0, int32_t cap, int32_t *__restrict__ count,
    int4 *__restrict__ list, int32_t n3) {
  const int64_t it = (int64_t)blockIdx.x * kListLaneThreads + threadIdx.x;
  if (it >= nclusters) return;
  const int64_t ci = seg_index(sg, it);
  if constexpr (MB) {
    if (box_active && box_active[cluster_box[ci]] == 0) return;
  }
  const double4 ai = c0[ci];
  const int64_t t = cluster_tower[ci];
  // an empty cluster of a region's tail, or one whose tower a partial build's register sorts left to its
  // repeat (no tower until then): an empty row
  if (ai.w == 0. || t < 0) {
    count[ci] = 0;
    return;
  }
  TowerGeom g = g1;
  int64_t tb = 0;
  if constexpr (MB) {
    const int32_t b = tbox[t];
    g = tgs[b];
    tb = tbase[b];
  }
  const int64_t tl = t - tb;
  const int tx = (int)(tl % g.nt[0]), ty = (int)(tl / g.nt[0]);
  double bi[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) bi[k] = bbox[6 * ci + k];
  const int side = 2 * g.R + 1, nvis = side * side;
  int4 *row = list + ci * (int64_t)cap;
  int32_t cnt = 0;
  for (int v = 0; v < nvis; ++v)
    cnt += visit_entries<true>(v, side, g, tb, tx, ty, bi, ai, cstart, ccount, bbox, c0, cap, cnt, row, ci, n3);
  count[ci] = cnt;
}

// The same rows with a group of kSplitLanes lanes per i-cluster: lane l of the group takes the visits
// l, l + kSplitLanes, ...; each lane counts its visit's entries, an inclusive scan over the group gives
// every visit its place in the row, and each lane then writes its visit's entries there.  The rows are
// those of the lane kernel entry by entry, while the dependent loads of one i-cluster's traversal (the
// bound of the lane kernel: ~850 per lane on the ensemble) are split over the group.
constexpr int kSplitLanes = 32;
constexpr int kListSplitThreads = 256;

template <bool MB>
__global__ __launch_bounds__(kListSplitThreads) void k_cluster_lists_split(
    int64_t nclusters, Segs sg, TowerGeom g1, const TowerGeom *__restrict__ tgs, const int64_t *__restrict__ tbase,
    const int32_t *__restrict__ tbox, const int32_t *__restrict__ cluster_box, const int32_t *__restrict__ box_active,
    const int32_t *__restrict__ cluster_tower, const int64_t *__restrict__ cstart, const int32_t *__restrict__ ccount,
    const double *__restrict__ bbox, const double4 *__restrict__ c0, int32_t cap, int32_t *__restrict__ count,
    int4 *__restrict__ list, int32_t n3) {
  const int sl = threadIdx.x & (kSplitLanes - 1);
  const int64_t it = ((int64_t)blockIdx.x * kListSplitThreads + threadIdx.x) / kSplitLanes;
  // every lane of a group holds the same i-cluster and the same decisions; no lane leaves before the scans
  const int64_t ci = it < nclusters ? seg_index(sg, it) : -1;
  bool act = ci >= 0;
  if constexpr (MB) act = act && (box_active == nullptr || box_active[cluster_box[ci]] != 0);
  double4 ai = make_double4(0., 0., 0., 0.);
  int64_t t = -1;
  if (act) {
    ai = c0[ci];
    t = cluster_tower[ci];
  }
  // an empty cluster of a region's tail, or one whose tower a partial build's register sorts left to its
  // repeat (no tower until then), gets an empty row
  const bool work = act && ai.w != 0. && t >= 0;
  TowerGeom g = g1;
  int64_t tb = 0;
  int tx = 0, ty = 0, side = 1, nvis = 0;
  double bi[6] = {0., 0., 0., 0., 0., 0.};
  if (work) {
    if constexpr (MB) {
      const int32_t b = tbox[t];
      g = tgs[b];
      tb = tbase[b];
    }
    const int64_t tl = t - tb;
    tx = (int)(tl % g.nt[0]);
    ty = (int)(tl / g.nt[0]);
    side = 2 * g.R + 1;
    nvis = side * side;
#pragma unroll
    for (int k = 0; k < 6; ++k) bi[k] = bbox[6 * ci + k];
  }
  // rounds of kSplitLanes visits; the count is made uniform over the wavefront for the scans
  int rounds = (nvis + kSplitLanes - 1) / kSplitLanes;
  for (int off = kSplitLanes; off < kWave; off <<= 1) {
    const int o = __shfl_xor(rounds, off, kWave);
    rounds = o > rounds ? o : rounds;
  }
  int4 *row = work ? list + ci * (int64_t)cap : nullptr;
  int32_t base = 0;
  for (int r = 0; r < rounds; ++r) {
    const int v = r * kSplitLanes + sl;
    const bool mine = work && v < nvis;
    const int32_t n =
        mine ? visit_entries<false>(v, side, g, tb, tx, ty, bi, ai, cstart, ccount, bbox, c0, cap, 0, row, ci, n3)
             : 0;
    int32_t inc = n;  // inclusive scan over the group
    for (int off = 1; off < kSplitLanes; off <<= 1) {
      const int32_t y = __shfl_up(inc, off, kSplitLanes);
      if (sl >= off) inc += y;
    }
    const int32_t total = __shfl(inc, kSplitLanes - 1, kSplitLanes);
    if (mine && n > 0)
      (void)visit_entries<true>(v, side, g, tb, tx, ty, bi, ai, cstart, ccount, bbox, c0, cap, base + inc - n, row,
                                ci, n3);
    base += total;
  }
  if (act && sl == 0) count[ci] = base;
}

// Second pass of the j-particle row build: one wavefront per i-cluster streams the members of its listed
// j-clusters (64 at a time) and keeps the real j within ilen of at least one real i-member, by the FP64
// length of the exact minimum-image fixed-point separation (the arithmetic of the Verlet-list build; it
// needs L > 2 ilen).  Survivors are compacted in stream order with a ballot and a prefix popcount, so the
// row is the union of the members' Verlet lists (members included), each j once.
// stats[2] is set unless every member and every row entry lies within L/2 - ilen of the cluster's first
// member (per dimension, minimum image), the member offsets included: |q| + |j| <= L/2 - ilen makes
// the separation q - j of the two offsets free of wrap-around, and the margin ilen covers the motion
// allowed between rebuilds (skin/2 per particle).  The force kernel then works in the first member's
// frame; otherwise it uses exact per-pair differences.
constexpr int kExpandThreads = 256;

// CUB (cubic boxes): the member test runs in FP32 on the separation in fixed-point units, against ilen^2
// in those units; tests whose FP32 d^2 lies in a band around it (cut_lo, cut_hi on the bit pattern; the
// FP32 error of d^2 is ~5 roundings, the band 2^-18 relative) are decided again by the FP64 test above,
// so the rows are the same.  Otherwise every test is the FP64 one.
// STAGED: each group first copies its concatenated sub-rows into LDS (ccap int32 per group, dynamic
// shared memory of the launch), so that entry e of the stream is one LDS read instead of a search over
// the sub-row offsets.
template <int M, int NS, int G, bool CUB, bool STAGED>
__global__ __launch_bounds__(kExpandThreads) void k_jrows_expand(
    int64_t nclusters, const int4 *__restrict__ posfx, const double4 *__restrict__ c0, int32_t ccap,
    const int32_t *__restrict__ ccount, const int32_t *__restrict__ jcl, FixGeom f, double ilen, uint32_t cut_lo,
    uint32_t cut_hi, int32_t cap, int32_t *__restrict__ count, int32_t *__restrict__ jrows,
    unsigned long long *__restrict__ stats) {
  constexpr int LG = kWave / G;  // lanes per i-cluster
  const int lane = threadIdx.x & (kWave - 1);
  const int grp = lane / LG, l = lane % LG;
  int64_t ic = (int64_t)(blockIdx.x * (kExpandThreads / kWave) + threadIdx.x / kWave) * G + grp;
  if constexpr (G == 1) ic = __builtin_amdgcn_readfirstlane((int)ic);
  const bool live = ic < nclusters;
  const int64_t icl = live ? ic : nclusters - 1;
  const double il2 = ilen * ilen;
  const int nri = (int)c0[icl].w;
  int4 xi[M];
#pragma unroll
  for (int q = 0; q < M; ++q) xi[q] = posfx[icl * M + q];
  // lanes whose cluster has a real member q
  unsigned long long real_m[M];
#pragma unroll
  for (int q = 0; q < M; ++q) real_m[q] = __builtin_amdgcn_ballot_w64(q < nri);
  // frame limits for the row entries: L/2 - ilen minus the largest member offset
  int64_t jlim[3];
  bool frame_bad = false;
  {
    int64_t qmax[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < M; ++q) {
      if (q < nri) {
        qmax[0] = max(qmax[0], llabs((int64_t)(int32_t)((uint32_t)xi[q].x - (uint32_t)xi[0].x)));
        qmax[1] = max(qmax[1], llabs((int64_t)(int32_t)((uint32_t)xi[q].y - (uint32_t)xi[0].y)));
        qmax[2] = max(qmax[2], llabs((int64_t)(int32_t)((uint32_t)xi[q].z - (uint32_t)xi[0].z)));
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      jlim[d] = (int64_t)(2147483648.0 - ilen * f.inv_scale[d]) - qmax[d];
      frame_bad = frame_bad || (live && jlim[d] < 0);
    }
  }
  // the NS sub-rows of the first pass, in order (an overflowed sub-row is rebuilt anyway)
  const int32_t subcap = ccap / NS;
  int pre[NS + 1];
  pre[0] = 0;
#pragma unroll
  for (int w = 0; w < NS; ++w) {
    const int32_t cw = ccount[icl * NS + w];
    pre[w + 1] = pre[w] + (cw < subcap ? cw : subcap);
  }
  const int nflat = live ? pre[NS] * M : 0;
  // the loop runs to the wave's longest stream so that the ballots stay uniform
  int nmax = nflat;
  if constexpr (G > 1) {
#pragma unroll
    for (int off = 32; off >= LG; off >>= 1) {
      const int o = __shfl_xor(nmax, off, kWave);
      nmax = o > nmax ? o : nmax;
    }
  }
  const int32_t *crow = jcl + icl * (int64_t)ccap;
  extern __shared__ int32_t s_entries[];
  int32_t *srow = STAGED ? s_entries + (int64_t)(threadIdx.x / LG) * ccap : nullptr;
  if constexpr (STAGED) {
    // the group's row in LDS (also for a dead group, whose icl is a real cluster: its lanes' clamped
    // prefetches below read real entries); a wavefront reads only its own groups' rows
#pragma unroll
    for (int w = 0; w < NS; ++w)
      for (int k = l; k < pre[w + 1] - pre[w]; k += LG) srow[pre[w] + k] = crow[w * subcap + k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // entry e of the concatenated sub-rows
  auto centry = [&](int e) {
    if constexpr (STAGED) return srow[e];
    int w = 0;
#pragma unroll
    for (int k = 1; k < NS; ++k) w += e >= pre[k] ? 1 : 0;
    int off = e;
#pragma unroll
    for (int k = 1; k < NS; ++k) off -= (w == k) ? pre[k] : 0;
    return crow[w * subcap + off];
  };
  int32_t *jrow = jrows + icl * (int64_t)cap;
  const unsigned long long gmask = (LG == kWave ? ~0ull : ((1ull << LG) - 1ull)) << (grp * LG);
  int32_t cnt = 0;
  // Every row holds at least one entry (the cluster's own, or a full sub-row), so with the element
  // index clamped to the stream's last one every lane, dead groups included, loads a real slot and the
  // prefetches need no guard.
  const int klast = pre[NS] * M - 1;
  int32_t jsn;
  int4 xjn;
  {
    const uint32_t k0 = (uint32_t)(l < klast ? l : klast);
    jsn = centry((int)(k0 / M)) * M + (int)(k0 % M);
    xjn = posfx[jsn];
  }
  for (int base = 0; base < nmax; base += LG) {
    const int32_t js = jsn;
    const int4 xj = xjn;
    const bool inrow = base + l < nflat;
    {
      const int kn = base + LG + l;
      const uint32_t kc = (uint32_t)(kn < klast ? kn : klast);
      jsn = centry((int)(kc / M)) * M + (int)(kc % M);
      xjn = posfx[jsn];
    }
    bool near = false;
    // the FP64 member test
    auto near64 = [&](int q, const int4 &pj) {
      const double dx = (double)(int32_t)((uint32_t)xi[q].x - (uint32_t)pj.x) * f.scale[0];
      const double dy = (double)(int32_t)((uint32_t)xi[q].y - (uint32_t)pj.y) * f.scale[1];
      const double dz = (double)(int32_t)((uint32_t)xi[q].z - (uint32_t)pj.z) * f.scale[2];
      return dx * dx + dy * dy + dz * dz <= il2;
    };
    if (inrow && xj.w >= 0) {  // padding slots carry type -1
      if constexpr (CUB) {
        // per member, the lanes inside and the lanes in the band, as wave masks (scalar unit)
        unsigned long long in_m = 0ull, band_m = 0ull;
#pragma unroll
        for (int q = 0; q < M; ++q) {
          const float dx = (float)(int32_t)((uint32_t)xi[q].x - (uint32_t)xj.x);
          const float dy = (float)(int32_t)((uint32_t)xi[q].y - (uint32_t)xj.y);
          const float dz = (float)(int32_t)((uint32_t)xi[q].z - (uint32_t)xj.z);
          const uint32_t ub = __float_as_uint(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)));
          const unsigned long long lo = __builtin_amdgcn_ballot_w64(ub < cut_lo) & real_m[q];
          in_m |= lo;
          band_m |= __builtin_amdgcn_ballot_w64(ub < cut_hi) & real_m[q] & ~lo;
        }
        near = (in_m >> lane) & 1ull;
        if (band_m != 0ull) {
          // rare; j passes through an empty asm so that the integer separations above are not kept
          int4 pj = xj;
          asm volatile("" : "+v"(pj.x), "+v"(pj.y), "+v"(pj.z));
          if ((band_m >> lane) & 1ull) {
#pragma unroll
            for (int q = 0; q < M; ++q) near = near || (q < nri && near64(q, pj));
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < M; ++q) near = near || (q < nri && near64(q, xj));
      }
      if (near) {
        frame_bad = frame_bad || llabs((int64_t)(int32_t)((uint32_t)xj.x - (uint32_t)xi[0].x)) > jlim[0] ||
                    llabs((int64_t)(int32_t)((uint32_t)xj.y - (uint32_t)xi[0].y)) > jlim[1] ||
                    llabs((int64_t)(int32_t)((uint32_t)xj.z - (uint32_t)xi[0].z)) > jlim[2];
      }
    }
    const unsigned long long ball = __ballot(near) & gmask;
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(ball >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ball, 0u));
    if (near && cnt + below < cap) jrow[cnt + below] = js;
    cnt += __popcll(ball);
  }
  if (live && l == 0) count[ic] = cnt;
  if (__ballot(frame_bad) != 0ull && lane == 0) stats[2] = 1ull;  // every writer stores the same value
}

__global__ void k_zero_u64(int64_t n, unsigned long long *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// FP32 position relative to the cluster reference point, particle type bits in w.  Padding slots sit at
// the sentinel +3e30 in every coordinate: any distance to them squares to +inf in FP32, which fails the
// cutoff test, while the difference itself stays finite (so masked terms contribute exact zeros).
// Written to both position buffers: the fused integrator of the FP32 cluster kernel writes the real
// slots of the second one each step, and the padding slots must hold the sentinel in both.  The FP64
// positions go to posd (both buffers too), and the largest |component| of the real slots' FP32 positions
// into the lines of ext.
__global__ void k_pos32(int64_t nslots, Segs sg, int32_t M, const double *__restrict__ x, const double *__restrict__ y,
                        const double *__restrict__ z, const int32_t *__restrict__ type,
                        const int32_t *__restrict__ orig, const double4 *__restrict__ c0, float4 *__restrict__ pos32,
                        float4 *__restrict__ pos32_next, double4 *__restrict__ posd, double4 *__restrict__ posd_next,
                        unsigned long long *__restrict__ ext, const int32_t *__restrict__ cluster_box,
                        const int32_t *__restrict__ box_active) {
  const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (it < nslots) {
    const int64_t s = seg_index(sg, it);
    if (!box_active || box_active[cluster_box[s / M]] != 0) {
      float4 p;
      double4 d;
      if (orig[s] < 0) {
        p = make_float4(kPadPos32, kPadPos32, kPadPos32, __int_as_float(0));
        d = make_double4(kDummyPos, kDummyPos, kDummyPos, 0.);
      } else {
        const double4 r = c0[s / M];
        d = make_double4(x[s], y[s], z[s], 0.);
        p = make_float4((float)(d.x - r.x), (float)(d.y - r.y), (float)(d.z - r.z), __int_as_float(type[s]));
        e = fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z)));
      }
      pos32[s] = p;
      pos32_next[s] = p;
      posd[s] = d;
      posd_next[s] = d;
    }
  }
  block_max_to_lines(e, ext);
}

void vcl_pos32(Ctx &c, const int32_t *active) {
  c.pos32.ensure(c.nslots);
  c.pos32_next.ensure(c.nslots);
  c.posd.ensure(c.nslots);
  c.posd_next.ensure(c.nslots);
  c.pos32_ext.ensure(kAccSlots * kAccStride);
  State &s = c.st[c.cur];
  const int64_t ni = active ? c.seg_nslots : c.nslots;
  const Segs sg = active ? c.seg_slots : Segs{nullptr, nullptr, 0};
  // a whole build starts the bound again; a partial one only raises it (the other boxes keep theirs)
  if (!active)
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(kAccSlots * kAccStride), 0, c.stream,
                       (int64_t)(kAccSlots * kAccStride), c.pos32_ext.p);
  if (ni == 0) return;
  hipLaunchKernelGGL(k_pos32, dim3(blocks_for(ni, 256)), dim3(256), 0, c.stream, ni, sg, c.M, s.x.p, s.y.p, s.z.p,
                     s.type.p, s.orig.p, c.c0.p, c.pos32.p, c.pos32_next.p, c.posd.p, c.posd_next.p, c.pos32_ext.p,
                     c.cluster_box.p, active);
  LJG_HIP(hipGetLastError());
}

// Member rows of the FP32 cluster kernel.  VCLClusterFunctor evaluates every member of every listed
// j-cluster against every i-member (VCLClusterFunctor.h:38-89); a listed j-cluster (bounding boxes within
// ilen) holds members farther than ilen from all i-members, which can never come within rc before the
// next build (each particle moves less than skin/2).  G i-clusters per wavefront stream their list's
// (entry, member) pairs and keep those within ilen of a real i-member, compacted in stream order with a
// ballot and a prefix popcount: xrow[ic * xstride + k] = {j slot, entry}, xcount[ic] of them.  The
// test is FP32 on the cluster-relative positions (the kernel's own arithmetic) against ilen^2 (1 + 2^-16),
// whose margin is far above the FP32 error of d^2, so a row holds every member within ilen in FP64 and
// at most a few more.  The cluster's own members are kept (the force kernel drops the pair of a member
// with itself).
constexpr int kXrowThreads = 256;
#ifndef LJG_XROWS_SMALL
#define LJG_XROWS_SMALL 32768
#endif

// A list entry and the FP32 position of one of its members, loaded together by the member-row pass.
struct EntryPos {
  int4 e;
  float4 p;
};
__device__ inline EntryPos make_pair_ef(const int4 &e, const float4 &p) { return EntryPos{e, p}; }
// The entry a lane of an empty row reads: cluster 0, member slots of a real cluster, never near anything
// it is tested against (the lane is past its row).
__device__ const int4 kZeroEntry = {0, 0, 0, 0};

template <int M, int G>
__global__ __launch_bounds__(kXrowThreads) void k_cluster_xrows(int64_t nitems, Segs sg,
                                                                const float4 *__restrict__ pos32, int32_t row,
                                                                const int32_t *__restrict__ lcount,
                                                                const int4 *__restrict__ list, float thr,
                                                                int32_t xstride, int32_t *__restrict__ xcount,
                                                                int2 *__restrict__ xrow) {
  constexpr int LG = kWave / G;
  const int lane = threadIdx.x & (kWave - 1);
  const int grp = lane / LG, l = lane % LG;
  const int64_t item = ((int64_t)blockIdx.x * (kXrowThreads / kWave) + threadIdx.x / kWave) * G + grp;
  const bool live = item < nitems;
  const int64_t ic = seg_index(sg, live ? item : nitems - 1);
  float xi[M], yi[M], zi[M];
#pragma unroll
  for (int q = 0; q < M; ++q) {
    const float4 p = pos32[ic * M + q];
    xi[q] = p.x;
    yi[q] = p.y;
    zi[q] = p.z;
  }
  const int nj = live ? lcount[ic] * M : 0;
  int nmax = nj;  // the loop runs to the wave's longest stream so that the ballots stay uniform
#pragma unroll
  for (int off = kWave / 2; off >= LG; off >>= 1) {
    const int o = __shfl_xor(nmax, off, kWave);
    nmax = o > nmax ? o : nmax;
  }
  const int4 *lr = list + ic * (int64_t)row;
  int2 *xr = xrow + ic * (int64_t)xstride;
  const unsigned long long gmask = (LG == kWave ? ~0ull : ((1ull << LG) - 1ull)) << (grp * LG);
  int32_t cnt = 0;
  // The entry and position of this lane's next candidate are loaded one step ahead, two steps per loop
  // trip so that the loaded values stay in their registers across the back edge.  A lane whose cluster
  // lists nothing (an empty cluster of a region's tail) reads a zero entry instead of its row, so every
  // load is in bounds without a guard.
  const uint32_t klast = nj > 0 ? (uint32_t)(nj - 1) : 0u;
  auto load_entry = [&](uint32_t kk) {
    const uint32_t kc = kk < klast ? kk : klast;
    const int4 *ep = nj > 0 ? lr + kc / M : &kZeroEntry;
    const int4 e = *ep;
    return make_pair_ef(e, pos32[(int64_t)(e.x & kClusterMask) * M + (kc % M)]);
  };
  auto x_step = [&](int base, const EntryPos &ep) {
    const int k = base + l;
    bool near = false;
    const int32_t js = (ep.e.x & kClusterMask) * M + (int32_t)((uint32_t)k % M);
    if (k < nj) {
      const float4 p = ep.p;
      if (p.x != kPadPos32) {
        const float jx = p.x + __int_as_float(ep.e.y), jy = p.y + __int_as_float(ep.e.z),
                    jz = p.z + __int_as_float(ep.e.w);
#pragma unroll
        for (int q = 0; q < M; ++q) {
          const float dx = xi[q] - jx, dy = yi[q] - jy, dz = zi[q] - jz;
          const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
          near = near || d2 <= thr;  // a padding i-member (at +3e30) is never near
        }
      }
    }
    const unsigned long long ball = __ballot(near) & gmask;
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(ball >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ball, 0u));
    if (near) xr[cnt + below] = make_int2(js, (int32_t)((uint32_t)k / M));
    cnt += __popcll(ball);
  };
  EntryPos pa = load_entry((uint32_t)l);
  for (int base = 0; base < nmax; base += 2 * LG) {
    const EntryPos pb = load_entry((uint32_t)(base + LG + l));
    x_step(base, pa);
    if (base + LG >= nmax) break;
    pa = load_entry((uint32_t)(base + 2 * LG + l));
    x_step(base + LG, pb);
  }
  if (live && l == 0) xcount[ic] = cnt;
}

// A wavefront per cluster of an unmarked box copies its member row (xcount[c] entries) to stride newstride.
__global__ __launch_bounds__(kXrowThreads) void k_restride_xrows(int64_t nclusters,
                                                                 const int32_t *__restrict__ cluster_box,
                                                                 const int32_t *__restrict__ box_active,
                                                                 const int32_t *__restrict__ count,
                                                                 const int2 *__restrict__ in, int32_t oldstride,
                                                                 int2 *__restrict__ out, int32_t newstride) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t c = (int64_t)blockIdx.x * (kXrowThreads / kWave) + threadIdx.x / kWave;
  if (c >= nclusters || box_active[cluster_box[c]] != 0) return;
  const int32_t n = count[c] < oldstride ? count[c] : oldstride;
  for (int32_t k = lane; k < n; k += kWave) out[c * (int64_t)newstride + k] = in[c * (int64_t)oldstride + k];
}

// The member rows after a cluster-entry build: every cluster's, or (partial ensemble build) the marked
// boxes' clusters only, the other boxes' rows moved to the new stride when the list rows grew.
void vcl_member_rows(Ctx &c, bool partial, const int32_t *active) {
  const int M = c.M;
  const int64_t ncl = c.nclusters;
  const int32_t xs = c.list_row * M;
  // A row holds the members within ilen at the box's last build: the unmarked boxes' rows must be kept
  // (rows made from later positions could miss pairs that come within rc before the box's next build).
  if (partial && !c.xrows_valid)
    throw HipError{hipErrorInvalidValue, "a partial build needs the member rows of the boxes it keeps", -2};
  const bool keep = partial;
  if (keep && c.xstride != xs) {
    c.xrow_tmp.ensure((size_t)ncl * xs + 1);
    constexpr int per_block = kXrowThreads / kWave;
    hipLaunchKernelGGL(k_restride_xrows, dim3(blocks_for(ncl, per_block)), dim3(kXrowThreads), 0, c.stream, ncl,
                       c.cluster_box.p, active, c.xcount.p, c.xrow.p, c.xstride, c.xrow_tmp.p, xs);
    LJG_HIP(hipGetLastError());
    std::swap(c.xrow, c.xrow_tmp);
  } else if (!keep) {
    c.xrow.ensure((size_t)ncl * xs + 1);
    c.xcount.ensure((size_t)ncl + 1);
  }
  c.xstride = xs;
  const Segs none{nullptr, nullptr, 0};
  const Segs sg = keep ? c.seg_clusters : none;
  const int64_t ni = keep ? c.seg_nclusters : ncl;
  const double ilen = c.prm.cutoff + c.prm.skin;
  const float thr = (float)(ilen * ilen * (1.0 + std::ldexp(1.0, -16)));
  if (ni > 0) {
    // a small build (few boxes of an ensemble) is bound by the latency of one wavefront's stream: a
    // wavefront per i-cluster then; four per wavefront otherwise
    const bool small = ni <= LJG_XROWS_SMALL;
    const int nb = blocks_for(ni, (kXrowThreads / kWave) * (small ? 1 : 4));
#define LJG_XROWS(MV, GV)                                                                                          \
  hipLaunchKernelGGL((k_cluster_xrows<MV, GV>), dim3(nb), dim3(kXrowThreads), 0, c.stream, ni, sg, c.pos32.p,       \
                     c.list_row, c.list_count.p, c.list.p, thr, xs, c.xcount.p, c.xrow.p)
#define LJG_XROWS_M(MV)                                                                                            \
  case MV:                                                                                                         \
    if (small) LJG_XROWS(MV, 1); else LJG_XROWS(MV, 4);                                                            \
    break;
    switch (M) {
      LJG_XROWS_M(1) LJG_XROWS_M(2) LJG_XROWS_M(4) LJG_XROWS_M(8)
      default: throw HipError{hipErrorInvalidValue, "cluster size must be 1, 2, 4 or 8", __LINE__};
    }
#undef LJG_XROWS_M
#undef LJG_XROWS
    LJG_HIP(hipGetLastError());
  }
  c.xrows_valid = true;
}

__global__ void k_fill_i32(int64_t n, int32_t *__restrict__ p, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_i32(Ctx &c, int32_t *p, int64_t n, int32_t v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, p, v);
  LJG_HIP(hipGetLastError());
}

// Average clusters per tower below which the cluster-entry lists use a lane (or a group of lanes) per
// i-cluster, and average particles per tower up to which the towers of at most 512 particles are sorted in
// registers.
constexpr int kLaneListsBelow = 24;
constexpr int kWaveSortBelow = 448;
// i-clusters below which those rows are made by a group of lanes per i-cluster: such a build (a few boxes
// of an ensemble, a small box) is bound by the latency of its dependent loads, ~850 per lane in the lane
// kernel, ~60 per lane in the group form; above it, by a lane each (the group form scans every visit
// twice, and a build this large fills the GPU with lanes anyway)
constexpr int64_t kSplitListsBelow = 1 << 18;

// The capacities a build keeps for the next one, with headroom for fluctuations (a melting lattice
// lengthens its rows by ~10% between rebuilds early on: by default a quarter of headroom keeps the two
// passes from running twice; ljg_params.row_headroom); a partial build keeps the stride the other boxes'
// rows are stored with.
static int32_t keep_capacity(const Ctx &c, bool partial, int32_t used, int32_t had) {
  if (partial) return had;
  if (c.prm.row_headroom < 0) return used > had ? used : (had > 2 * used ? used : had);  // none
  const int32_t pct = c.prm.row_headroom == 0 ? 25 : c.prm.row_headroom;
  const int32_t want = used + (int32_t)((int64_t)used * pct / 100) + 8;
  return want > had ? want : (had > 2 * want ? want : had);
}

bool vcl_finish_lists(Ctx &c) {
  if (!c.lists_pending) return true;
  c.lists_pending = false;
  unsigned long long st[6] = {0, 0, 0, 0, 0, 0};
  stream_wait(c, c.list_stats.p, 6, st);
  const int32_t maxc = (int32_t)st[1], maxj = (int32_t)st[4];
  const bool overflow = st[5] != 0ull || maxc > c.pend_cap / kJclSplit || maxj > c.pend_jcap;
  c.jrows_frame_ok = st[2] == 0ull;
  if (overflow) {
    c.jrows_valid = false;  // the rows are incomplete: the next use builds again
    return false;
  }
  c.jccap = kJclSplit * keep_capacity(c, false, maxc, c.pend_cap / kJclSplit);
  c.jcap = keep_capacity(c, false, maxj, c.pend_jcap);
  c.nentries = (int64_t)st[3];
  return !(c.pend_fast && !c.jrows_frame_ok);
}

void vcl_build(Ctx &c) {
  c.lists_pending = false;
  const int M = c.M;
  const double ilen = c.prm.cutoff + c.prm.skin;
  // Tower grid: ClusterTowerBlock2D::estimateOptimalGridSideLength (ClusterTowerBlock2D.h:140-168) with
  // the owned towers only (periodic images come from shift vectors) and resize (:89-114); one per box of
  // an ensemble, numbered box by box.
  const bool ens = c.nboxes > 0;
  if (ens) {
    ensemble_geometry(c);
  } else {
    c.tg = make_tower_geom(c.prm.box_min, c.prm.box_max, c.n, M, ilen);
  }
  const TowerGeom &g = c.tg;
  const int64_t ntow = ens ? c.box_ntowers : (int64_t)g.nt[0] * g.nt[1];
  c.ntowers = ntow;
  // An ensemble laid out in its regions may rebuild only the boxes box_active marks, in place: their
  // particles stay in their own regions and every other box keeps its slots, clusters and lists.
  const bool partial = ens && c.build_partial && c.regions_valid;
  const int32_t *active = partial ? c.seg_active : nullptr;

  const int64_t ns = c.nslots;
  State &s = c.st[c.cur];
  c.tower_of.ensure(ns);
  c.rank.ensure(ns);
  c.tower_count.ensure(ntow + 1);
  c.tower_start.ensure(ntow + 1);
  c.tower_items.ensure(c.n + 1);
  c.tower_z.ensure(c.n + 1);
  c.tower_id.ensure(c.n + 1);
  c.ctow_count.ensure(ntow + 1);
  c.ctow_start.ensure(ntow + 1);
  c.flag.ensure(4);
  fill_i32(c, c.tower_count.p, ntow + 1, 0);
  c.tstats.ensure(4);
  if (ens) {
    // per-box layouts inside the fixed regions; only the longest tower is needed from the stats
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c.stream, (int64_t)4, c.tstats.p);
    LJG_HIP(hipGetLastError());
    ensemble_tower_hist(c, active);
    ensemble_layout(c, active);
  } else {
    hipLaunchKernelGGL(k_tower_hist, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, s.x.p, s.y.p, s.z.p,
                       s.orig.p, g, wrap_args(c.prm), c.tower_of.p, c.rank.p, c.tower_count.p);
    LJG_HIP(hipGetLastError());
  }
  // tower offsets, clusters per tower and their offsets, and the sizes for the one synchronisation
  if (ens) {
  } else if (ntow <= kLayoutOneBlockMax) {
    hipLaunchKernelGGL(k_tower_layout, dim3(1), dim3(kLayoutThreads), 0, c.stream, ntow, c.tower_count.p, M,
                       c.tower_start.p, c.ctow_count.p, c.ctow_start.p, c.tstats.p);
    LJG_HIP(hipGetLastError());
  } else {
    const int64_t nblk = (ntow + kLayoutChunk - 1) / kLayoutChunk;
    c.scan_tmp.ensure(2 * (size_t)nblk + (size_t)nblk / 2 + 1);
    int64_t *bsum = c.scan_tmp.p;
    int32_t *bmax = reinterpret_cast<int32_t *>(c.scan_tmp.p + 2 * nblk);
    hipLaunchKernelGGL(k_layout_sums, dim3((unsigned)nblk), dim3(kLayoutBlock), 0, c.stream, ntow, c.tower_count.p, M,
                       bsum, bmax);
    LJG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_layout_top, dim3(1), dim3(kLayoutThreads), 0, c.stream, nblk, ntow, bsum, bmax,
                       c.tower_start.p, c.ctow_start.p, c.tstats.p);
    LJG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_layout_apply, dim3((unsigned)nblk), dim3(kLayoutBlock), 0, c.stream, ntow, c.tower_count.p,
                       M, bsum, c.tower_start.p, c.ctow_count.p, c.ctow_start.p);
    LJG_HIP(hipGetLastError());
  }
  // a partial build launches over the marked boxes' segments only
  const Segs none{nullptr, nullptr, 0};
  const Segs sg_slots = partial ? c.seg_slots : none, sg_towers = partial ? c.seg_towers : none;
  const Segs sg_clusters = partial ? c.seg_clusters : none;
  const int64_t n_slots = partial ? c.seg_nslots : ns, n_towers = partial ? c.seg_ntowers : ntow;
  const bool guess = partial && !c.partial_tower_sync;
  // The sizes' marker goes before the scatter and the clearing of the list statistics, which need no
  // size: they run while the host waits for it, and the first launch after the wait is a long one.
  const unsigned long long tmark = guess ? 0ull : stream_mark(c, c.tstats.p, 4);
  // list_stats: [0], [1] total and longest row of the cluster-pair pass, [2] frame flag of the expansion
  // pass, [3], [4] total and longest j-particle row, [5] sort tile exceeded; read back once at the end
  c.list_stats.ensure(6);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c.stream, (int64_t)6, c.list_stats.p);
  LJG_HIP(hipGetLastError());
  if (n_slots > 0) {
    hipLaunchKernelGGL(k_tower_scatter, dim3(blocks_for(n_slots, 256)), dim3(256), 0, c.stream, n_slots, sg_slots,
                       c.tower_of.p, c.rank.p, c.tower_start.p, s.z.p, s.id.p, c.tower_items.p, c.tower_z.p,
                       c.tower_id.p);
    LJG_HIP(hipGetLastError());
  }
  // The sizes: the longest tower picks the sort instances (only those are launched: an instance with a
  // large LDS tile or many registers waits for room on a GPU busy with other work).  A partial build uses
  // the expected longest tower of the marked boxes instead of waiting for the readback, and launches the
  // register sorts up to it: a longer tower (rare) is reported by the longest instance launched, the
  // permutation then skips, and the build runs again with the readback.
  unsigned long long ts[4] = {0, 0, 0, 0};
  if (guess) {
    ts[1] = (unsigned long long)(c.seg_max_count < 8 * kWave ? c.seg_max_count : 8 * kWave);
  } else {
    stream_spin(c, tmark, 4, ts);
  }
  const int64_t ncl = ens ? c.box_clusters : (int64_t)ts[2];
  if (ts[1] > (unsigned long long)kTowerSortCap)
    throw HipError{hipErrorInvalidValue, "tower exceeds the 4096-particle sort tile", -1};
  c.nclusters = ncl;
  const int64_t new_slots = ncl * M;
  c.perm.ensure(new_slots);
  c.cluster_tower.ensure(ncl);
  if (ens) ensemble_region_prep(c, active);  // the regions' tails stay padding
  // towers of up to 512 particles sort in registers, a wavefront each (three instances by length, each
  // launched only if some tower is in its range); longer ones in LDS
  const bool wave_sort = ens || (int64_t)ts[0] <= (int64_t)kWaveSortBelow * ntow;
  const int32_t min_n = wave_sort ? 8 * kWave : 0;
  if (wave_sort) {
    constexpr int per_block = kSortWaveThreads / kWave;
#define LJG_WAVE_SORT(E, OVER)                                                                                  \
  if (n_towers > 0) {                                                                                           \
    hipLaunchKernelGGL(k_tower_sort_wave<E>, dim3(blocks_for(n_towers, per_block)), dim3(kSortWaveThreads), 0,  \
                       c.stream, n_towers, sg_towers, c.tower_start.p, c.tower_count.p, c.tower_items.p,        \
                       c.tower_z.p, c.tower_id.p, c.ctow_start.p, c.ctow_count.p, M, c.perm.p, c.cluster_tower.p, \
                       OVER);                                                                                   \
    LJG_HIP(hipGetLastError());                                                                                 \
  }
    // every tower is sorted by exactly one instance; the longest tower (ts[1]) and the shortest non-empty
    // one (ts[3], 0 when not known) bound which ones can have work
    unsigned long long *over = guess ? c.list_stats.p + 5 : nullptr;
    const int top = ts[1] > (unsigned long long)(4 * kWave) ? 8 : ts[1] > (unsigned long long)kWave ? 4 : 1;
    const unsigned long long shortest = ens ? 0ull : ts[3];
#define LJG_PAIR_SORT(E, OVER)                                                                                  \
  if (n_towers > 0) {                                                                                           \
    hipLaunchKernelGGL(k_tower_sort_pair<E>, dim3((unsigned)n_towers), dim3(kSortPairThreads), 0, c.stream,    \
                       n_towers, sg_towers, c.tower_start.p, c.tower_count.p, c.tower_items.p, c.tower_z.p,     \
                       c.tower_id.p, c.ctow_start.p, c.ctow_count.p, M, c.perm.p, c.cluster_tower.p, OVER);     \
    LJG_HIP(hipGetLastError());                                                                                 \
  }
    if (shortest <= (unsigned long long)kWave) LJG_WAVE_SORT(1, top == 1 ? over : nullptr)
    if (top >= 4 && shortest <= (unsigned long long)(4 * kWave)) LJG_PAIR_SORT(2, top == 4 ? over : nullptr)
    if (top == 8) LJG_PAIR_SORT(4, over)
#undef LJG_PAIR_SORT
#undef LJG_WAVE_SORT
  }
  if (ts[1] > (unsigned long long)min_n && n_towers > 0) {
    hipLaunchKernelGGL(k_tower_sort<kSortSmall>, dim3((unsigned)n_towers), dim3(kSortThreads), 0, c.stream, n_towers,
                       sg_towers, c.tower_start.p, c.tower_count.p, c.tower_items.p, s.z.p, s.id.p, c.ctow_start.p,
                       c.ctow_count.p, M, c.perm.p, c.cluster_tower.p, c.list_stats.p + 5, min_n);
    LJG_HIP(hipGetLastError());
  }
  if (ts[1] > (unsigned long long)kSortSmall && n_towers > 0) {
    hipLaunchKernelGGL(k_tower_sort<kTowerSortCap>, dim3((unsigned)n_towers), dim3(kSortThreads), 0, c.stream,
                       n_towers, sg_towers, c.tower_start.p, c.tower_count.p, c.tower_items.p, s.z.p, s.id.p,
                       c.ctow_start.p, c.ctow_count.p, M, c.perm.p, c.cluster_tower.p, c.list_stats.p + 5, min_n);
    LJG_HIP(hipGetLastError());
  }
  const bool jrows = c.want_jrows;
  if (jrows && ens) throw HipError{hipErrorInvalidValue, "j-particle rows on an ensemble of boxes", -2};
  if (jrows) {
    // fixed-point positions of every slot (padding: type -1), written by the permutation
    verlet_geometry(c);
    c.posfx.ensure(new_slots);
  }
  if (partial) permute_active(c, c.perm.p, active, guess ? c.list_stats.p + 5 : nullptr);
  else permute_state(c, c.perm.p, new_slots, jrows ? c.posfx.p : nullptr);
  c.regions_valid = ens;
  c.M = M;
  State &t = c.st[c.cur];
  c.bbox.ensure(6 * ncl);
  c.c0.ensure(ncl);
  c.list_count.ensure(ncl * kJclSplit + 1);
  const int64_t n_bbox = partial ? c.seg_nslots : new_slots;
  if (n_bbox > 0) {
    hipLaunchKernelGGL(k_cluster_bbox, dim3(blocks_for(n_bbox, 256)), dim3(256), 0, c.stream, n_bbox, sg_slots, M,
                       t.x.p, t.y.p, t.z.p, t.orig.p, c.bbox.p, c.c0.p,
                       ens ? c.cluster_box.p : (const int32_t *)nullptr, active,
                       ens ? c.list_count.p : (int32_t *)nullptr);
    LJG_HIP(hipGetLastError());
  }
  if (jrows && M > kMaxM) throw HipError{hipErrorInvalidValue, "cluster size above 8", -2};
  if (jrows) c.jcount.ensure(ncl + 1);
  // Both passes are queued back to back and checked with one synchronisation; a row that overflowed
  // its capacity (the expansion pass reads at most ccap entries) makes both run again with larger rows.
  // cluster-entry lists: a lane per i-cluster when the towers hold few clusters, unless the caller chose
  // (list_mode 3: the lane rows with a group of lanes per i-cluster; the automatic choice takes it for the
  // lane rows of builds of fewer than kSplitListsBelow i-clusters)
  const bool auto_lane = c.prm.list_mode == 0 && ncl < (int64_t)kLaneListsBelow * ntow;
  const bool split_lists =
      c.prm.list_mode == 3 || (auto_lane && (partial ? c.seg_nclusters : ncl) < kSplitListsBelow);
  const bool lane_lists = c.prm.list_mode == 2 || auto_lane || split_lists;
  // (a partial build starts from the stride the other boxes' rows are stored with)
  int32_t cap = jrows ? (c.jccap > 0 ? c.jccap : kJclSplit * 24)
                      : partial ? c.list_row : (c.list_cap > 0 ? c.list_cap : 64);
  int32_t jcap = c.jcap > 0 ? c.jcap : 160;
  bool pos32_queued = false;
  for (;;) {
    if (jrows) {
      c.jcl.ensure((size_t)ncl * cap + 1);
      c.jlist.ensure((size_t)ncl * jcap + 1);
    } else {
      c.list.ensure((size_t)ncl * cap + 1);
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c.stream, (int64_t)5, c.list_stats.p);
    LJG_HIP(hipGetLastError());
    const TowerGeom *no_tg = nullptr;
    const int64_t *no_i64 = nullptr;
    const int32_t *no_i32 = nullptr;
    const int64_t n_lane = partial ? c.seg_nclusters : ncl;
    const int32_t n3 = c.want_n3 && !jrows ? 1 : 0;
    if (n_towers == 0 || n_lane == 0)
      ;  // no box marked
    else if (jrows)
      hipLaunchKernelGGL((k_cluster_lists_tower<true, kJclSplit, false>), dim3((unsigned)ntow),
                         dim3(kListThreads * kJclSplit), 0, c.stream, ntow, none, g, no_tg, no_i64, no_i32, no_i32,
                         c.ctow_start.p, c.ctow_count.p, c.bbox.p, c.c0.p, cap, c.list_count.p, (int4 *)nullptr,
                         c.jcl.p, (int32_t)0);
    else if (split_lists && ens)
      hipLaunchKernelGGL((k_cluster_lists_split<true>), dim3(blocks_for(n_lane * kSplitLanes, kListSplitThreads)),
                         dim3(kListSplitThreads), 0, c.stream, n_lane, sg_clusters, g, c.box_tg.p, c.box_tbase.p,
                         c.tower_box.p, c.cluster_box.p, active, c.cluster_tower.p, c.ctow_start.p, c.ctow_count.p,
                         c.bbox.p, c.c0.p, cap, c.list_count.p, c.list.p, n3);
    else if (split_lists)
      hipLaunchKernelGGL((k_cluster_lists_split<false>), dim3(blocks_for(ncl * kSplitLanes, kListSplitThreads)),
                         dim3(kListSplitThreads), 0, c.stream, ncl, none, g, no_tg, no_i64, no_i32, no_i32, no_i32,
                         c.cluster_tower.p, c.ctow_start.p, c.ctow_count.p, c.bbox.p, c.c0.p, cap, c.list_count.p,
                         c.list.p, n3);
    else if (lane_lists && ens)
      hipLaunchKernelGGL((k_cluster_lists_lane<true>), dim3(blocks_for(n_lane, kListLaneThreads)),
                         dim3(kListLaneThreads), 0, c.stream, n_lane, sg_clusters, g, c.box_tg.p, c.box_tbase.p, c.tower_box.p, c.cluster_box.p, active,
                         c.cluster_tower.p, c.ctow_start.p, c.ctow_count.p, c.bbox.p, c.c0.p, cap, c.list_count.p,
                         c.list.p, n3);
    else if (lane_lists)
      hipLaunchKernelGGL((k_cluster_lists_lane<false>), dim3(blocks_for(ncl, kListLaneThreads)),
                         dim3(kListLaneThreads), 0, c.stream, ncl, none, g, no_tg, no_i64, no_i32, no_i32, no_i32,
                         c.cluster_tower.p, c.ctow_start.p, c.ctow_count.p, c.bbox.p, c.c0.p, cap, c.list_count.p,
                         c.list.p, n3);
    else if (ens)
      hipLaunchKernelGGL((k_cluster_lists_tower<false, 1, true>), dim3((unsigned)n_towers), dim3(kListThreads), 0,
                         c.stream, n_towers, sg_towers, g, c.box_tg.p, c.box_tbase.p, c.tower_box.p, active, c.ctow_start.p, c.ctow_count.p,
                         c.bbox.p, c.c0.p, cap, c.list_count.p, c.list.p, (int32_t *)nullptr, n3);
    else
      hipLaunchKernelGGL((k_cluster_lists_tower<false, 1, false>), dim3((unsigned)ntow), dim3(kListThreads), 0, c.stream,
                         ntow, none, g, no_tg, no_i64, no_i32, no_i32, c.ctow_start.p, c.ctow_count.p, c.bbox.p, c.c0.p,
                         cap, c.list_count.p, c.list.p, (int32_t *)nullptr, n3);
    LJG_HIP(hipGetLastError());
    count_stats(c, c.list_count.p, jrows ? ncl * kJclSplit : ncl, c.list_stats.p);
    if (jrows) {
      // i-clusters per wavefront of the expansion pass (tuned on the 10^6-particle liquid: 4 for M >= 4)
      // (one i-cluster per wave below 2048 clusters: latency-bound small builds)
      const int eg = c.prm.build_groups > 0 ? c.prm.build_groups
                                             : (M >= 4 ? (ncl <= 2048 ? 1 : 4) : M == 2 ? 4 : 8);
      // FP32 member test of the cubic variant: ilen^2 in fixed-point units, band 2^-18 relative
      uint32_t cut_lo = 0, cut_hi = 0;
      cutoff_band((float)(ilen * ilen * c.fg.inv_scale[0] * c.fg.inv_scale[0]), std::ldexp(1.0, -18), cut_lo, cut_hi);
// the staged form when the block's rows fit in 64 KiB of LDS (4 G rows of cap int32)
#define LJG_EXPAND_LAUNCH(MV, GV, CUBV)                                                                           \
  {                                                                                                               \
    const size_t lds = (size_t)(kExpandThreads / (kWave / GV)) * (size_t)cap * sizeof(int32_t);                   \
    if (lds <= 65536)                                                                                             \
      hipLaunchKernelGGL((k_jrows_expand<MV, kJclSplit, GV, CUBV, true>), dim3(nb), dim3(kExpandThreads),         \
                         (unsigned)lds, c.stream, ncl, c.posfx.p, c.c0.p, cap, c.list_count.p, c.jcl.p, c.fg,     \
                         ilen, cut_lo, cut_hi, jcap, c.jcount.p, c.jlist.p, c.list_stats.p);                      \
    else                                                                                                          \
      hipLaunchKernelGGL((k_jrows_expand<MV, kJclSplit, GV, CUBV, false>), dim3(nb), dim3(kExpandThreads), 0,     \
                         c.stream, ncl, c.posfx.p, c.c0.p, cap, c.list_count.p, c.jcl.p, c.fg, ilen, cut_lo,      \
                         cut_hi, jcap, c.jcount.p, c.jlist.p, c.list_stats.p);                                    \
  }
#define LJG_EXPAND(MV, GV)                                                                                        \
  if (M == MV && eg == GV) {                                                                                      \
    const int nb = blocks_for(ncl, (kExpandThreads / kWave) * GV);                                               \
    if (c.fg.cubic)                                                                                               \
      LJG_EXPAND_LAUNCH(MV, GV, true)                                                                             \
    else                                                                                                          \
      LJG_EXPAND_LAUNCH(MV, GV, false)                                                                            \
    goto expanded;                                                                                                \
  }
      LJG_EXPAND(1, 4) LJG_EXPAND(1, 8) LJG_EXPAND(2, 2) LJG_EXPAND(2, 4) LJG_EXPAND(4, 1) LJG_EXPAND(4, 2)
      LJG_EXPAND(4, 4) LJG_EXPAND(8, 1) LJG_EXPAND(8, 2) LJG_EXPAND(8, 4)
#undef LJG_EXPAND
#undef LJG_EXPAND_LAUNCH
      throw HipError{hipErrorInvalidValue, "unsupported (cluster size, expansion groups) combination", -2};
    expanded:
      LJG_HIP(hipGetLastError());
      count_stats(c, c.jcount.p, ncl, c.list_stats.p + 3);
    }
    // A j-row build the step loop follows with its force at once: the lengths are read after that launch,
    // which exits on the device if a row overflowed (vcl_finish_lists).  The first build of a context
    // (capacities not yet measured) is checked here.
    if (jrows && c.defer_lists && !partial && !ens && c.jccap > 0 && c.jcap > 0) {
      c.lists_pending = true;
      c.pend_cap = cap;
      c.pend_jcap = jcap;
      c.list_row = jcap;
      c.jrows_valid = true;
      c.lists_n3 = false;
      c.built_M = M;
      c.layout = Layout::Clusters;
      c.xrows_valid = false;
      return;
    }
    // the FP32 positions of the cluster kernels need no list: queued before the wait, they are made while
    // the host waits (once, also when the lists are made again with larger rows; not in a partial build
    // on guessed sizes, which may find that nothing was permuted and start again)
    if (!jrows && !guess && !pos32_queued) {
      vcl_pos32(c, active);
      pos32_queued = true;
    }
    unsigned long long st[6] = {0, 0, 0, 0, 0, 0};
    stream_wait(c, c.list_stats.p, 6, st);
    if (st[5] && guess) {
      // a marked tower was longer than the register sorts take: nothing was permuted, so build again
      // with the longest tower read back first
      c.partial_tower_sync = true;
      try {
        vcl_build(c);
      } catch (...) {
        c.partial_tower_sync = false;
        throw;
      }
      c.partial_tower_sync = false;
      return;
    }
    if (st[5]) throw HipError{hipErrorInvalidValue, "tower exceeds the 4096-particle sort tile", -1};
    const int32_t maxc = (int32_t)st[1], maxj = (int32_t)st[4];
    const int32_t split = jrows ? kJclSplit : 1;  // first-pass rows hold `split` sub-rows of cap/split
    const bool ok1 = maxc <= cap / split, ok2 = !jrows || maxj <= jcap;
    if (!ok1) {
      const int32_t grown = split * (maxc + maxc / 8 + 8);
      // the rows a partial build keeps move to the longer stride; the marked boxes' rows are made again
      if (partial) ensemble_restride_lists(c, active, cap, grown);
      cap = grown;
    }
    if (jrows && ok1 && !ok2) jcap = maxj + maxj / 8 + 8;
    if (!(ok1 && ok2)) continue;
    // keep the row capacities for the next build
    auto keep = [&c, partial](int32_t used, int32_t had) { return keep_capacity(c, partial, used, had); };
    if (jrows) {
      c.jccap = kJclSplit * keep(maxc, cap / kJclSplit);
      c.jcap = keep(maxj, jcap);
      c.nentries = (int64_t)st[3];
      c.jrows_frame_ok = st[2] == 0;
      c.list_row = jcap;
    } else {
      c.list_cap = keep(maxc, cap);
      c.nentries = (int64_t)st[0];
      c.list_row = cap;
    }
    break;
  }
  c.jrows_valid = jrows;
  c.lists_n3 = c.want_n3 && !jrows;
  c.built_M = M;
  c.layout = Layout::Clusters;
  if (!jrows && !pos32_queued) vcl_pos32(c, active);
  if (!jrows && !c.lists_n3 && !c.prm.whole_cluster_rows) vcl_member_rows(c, partial, active);
  else c.xrows_valid = false;
}

}  // namespace ljg
