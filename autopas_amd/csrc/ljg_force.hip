// ljg_force.hip -- Lennard-Jones 12-6 force kernels for gfx950 (wave64).
//
//  * k_force_lc_fp64: the plain FP64 kernel.  One lane per particle (cell order), full shell over the 27
//    neighbour cells of CellBlock3D with periodic image shifts, FP64 pair term written as LJFunctor.h:
//    288-320, FP64 sums.  No atomics, no Newton-3.
//  * k_force_cluster<M, FP32>: cluster-pair tiles.  One wave per i-cluster of M particles; the i-cluster
//    (positions, types) is wave-uniform and lives in scalar registers, the 64 lanes stream the j-particles
//    of the i-cluster's neighbour-cluster list (M consecutive lanes hold one j-cluster), and every lane
//    evaluates its j against all M i-particles.  Per-lane partial forces are reduced across the wave with
//    a reduce-scatter butterfly (each step halves the values a lane carries), in FP64.
//    FP32 variant: positions are FP32 offsets from each cluster's reference point c0 (|offset| ~ 1 sigma);
//    the per-entry offset c0_j + shift - c0_i is formed in FP64 and rounded once, so pair distances carry
//    absolute errors ~1e-7 sigma instead of ~1e-5 for raw FP32 coordinates in a 100-sigma box.  The LJ
//    term runs in FP32 (hardware reciprocal), per-lane partial sums in FP32 over the lane's few j, and the
//    cross-lane and final sums in FP64.
//    FP64 variant: same tiles, absolute FP64 positions, pair term exactly as LJFunctor.h:288-320.
// Each kernel also counts the pairs with r^2 <= rc^2 (ballot + popcount on the scalar unit) and, when
// asked, the potential energy (6x, shift included) and virial per LJFunctor.h:174-197 / :322-345.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "ljg_internal.h"

namespace ljg {

struct ForceArgs {
  double L[3];
  double rc2;
  int32_t ntypes;
  double eps24, sig2, shift6;  // single-type constants
};

constexpr int kFThreads = 256;
constexpr int kWavesPerBlock = kFThreads / kWave;

// ---- block-level deterministic partials ------------------------------------------------------------
// Per-block partial sums (fixed tree, so reproducible); the block's pair count is also added to acc's
// per-XCD lines when acc is set (integer addition, so the order of the atomics does not matter).
__device__ inline void block_partials(double e, double v, int64_t h, double *pe, double *pv, int64_t *ph,
                                      bool globals, unsigned long long *acc) {
  __shared__ double se[kFThreads], sv[kFThreads];
  __shared__ int64_t sh[kFThreads];
  se[threadIdx.x] = globals ? e : 0.;
  sv[threadIdx.x] = globals ? v : 0.;
  sh[threadIdx.x] = h;
  __syncthreads();
  for (int s = kFThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sv[threadIdx.x] += sv[threadIdx.x + s];
      sh[threadIdx.x] += sh[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pe[blockIdx.x] = se[0];
    pv[blockIdx.x] = sv[0];
    ph[blockIdx.x] = sh[0];
    if (acc) acc_add(acc, (unsigned long long)sh[0]);
  }
}

// ---- plain FP64 linked-cells kernel ------------------------------------------------------------------
template <bool MIX>
__global__ __launch_bounds__(kFThreads) void k_force_lc_fp64(
    int64_t n, const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const int32_t *__restrict__ type, const int32_t *__restrict__ cell_of, const int64_t *__restrict__ cell_start,
    CellGeom g, ForceArgs a, const double *__restrict__ tab, double *__restrict__ fx, double *__restrict__ fy,
    double *__restrict__ fz, double *__restrict__ pe, double *__restrict__ pv, int64_t *__restrict__ ph,
    int32_t globals, unsigned long long *__restrict__ acc, SkipIf sk, Owned own) {
  if (skip_launch(sk)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double ax = 0., ay = 0., az = 0., ae = 0., av = 0.;
  int64_t hits = 0;
  if (i < n) {
    const double xi = x[i], yi = y[i], zi = z[i];
    const int ti = MIX ? type[i] : 0;
    const int32_t c = cell_of[i];
    const int cx = c % g.nh[0], cy = (c / g.nh[0]) % g.nh[1], cz = c / (g.nh[0] * g.nh[1]);
    const int no0 = g.nh[0] - 2, no1 = g.nh[1] - 2, no2 = g.nh[2] - 2;
    for (int oz = -1; oz <= 1; ++oz) {
      int jz = cz - 1 + oz;
      double shz = 0.;
      if (jz < 0) { jz += no2; shz = -g.L[2]; } else if (jz >= no2) { jz -= no2; shz = g.L[2]; }
      for (int oy = -1; oy <= 1; ++oy) {
        int jy = cy - 1 + oy;
        double shy = 0.;
        if (jy < 0) { jy += no1; shy = -g.L[1]; } else if (jy >= no1) { jy -= no1; shy = g.L[1]; }
        for (int ox = -1; ox <= 1; ++ox) {
          int jx = cx - 1 + ox;
          double shx = 0.;
          if (jx < 0) { jx += no0; shx = -g.L[0]; } else if (jx >= no0) { jx -= no0; shx = g.L[0]; }
          const int64_t cidx = ((int64_t)(jz + 1) * g.nh[1] + (jy + 1)) * g.nh[0] + (jx + 1);
          const int64_t sb = cell_start[cidx], se = cell_start[cidx + 1];
          for (int64_t j = sb; j < se; ++j) {
            const double drx = xi - (x[j] + shx);
            const double dry = yi - (y[j] + shy);
            const double drz = zi - (z[j] + shz);
            const double dr2 = drx * drx + dry * dry + drz * drz;
            double eps24 = a.eps24, sig2 = a.sig2, shift6 = a.shift6;
            if (MIX) {
              const int k = ti * a.ntypes + type[j];
              eps24 = tab[3 * k];
              sig2 = tab[3 * k + 1];
              shift6 = tab[3 * k + 2];
            }
            const bool mask = dr2 <= a.rc2 && dr2 > 0.;
            const double invdr2 = 1. / dr2;
            const double lj2 = sig2 * invdr2;
            const double lj6 = lj2 * lj2 * lj2;
            const double lj12 = lj6 * lj6;
            const double lj12m6 = lj12 - lj6;
            const double fac = mask ? eps24 * (lj12 + lj12m6) * invdr2 : 0.;
            ax += drx * fac;
            ay += dry * fac;
            az += drz * fac;
            if (globals) {
              ae += mask ? eps24 * lj12m6 + shift6 : 0.;
              av += drx * (drx * fac) + dry * (dry * fac) + drz * (drz * fac);
            }
            hits += mask ? 1 : 0;
          }
        }
      }
    }
    fx[i] = ax;
    fy[i] = ay;
    fz[i] = az;
    if (globals && !is_owned(own, i)) ae = av = 0.;
  }
  block_partials(ae, av, hits, pe, pv, ph, globals != 0, acc);
}

// ---- cluster-pair tiles -------------------------------------------------------------------------------
constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }
constexpr int pow2_at_least(int x) { return x <= 1 ? 1 : 2 * pow2_at_least((x + 1) / 2); }

// Reduce-scatter of v[NV] over a group of LG consecutive lanes (LG, NV powers of two).  Each butterfly
// step halves the values a lane carries; once a lane carries one value the remaining steps add pairs.
// Afterwards lane l holds the group totals of values [q*R, q*R+R) in v[0..R-1], R = NV >> min(log LG,
// log NV), q = group_value_slot<NV, LG>(l).
template <int NV, int LG>
__device__ inline void group_reduce_scatter(double (&v)[NV], int l) {
  constexpr int KL = ilog2(LG), KV = ilog2(NV), K = KL < KV ? KL : KV;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int step = LG >> (k + 1);
    const int half = NV >> (k + 1);
    const bool upper = (l & step) != 0;
#pragma unroll
    for (int q = 0; q < half; ++q) {
      // select values, not addresses: a select of two array slots keeps the array in scratch memory
      const double lo = v[q], hi = v[q + half];
      const double send = upper ? lo : hi;
      const double keep = upper ? hi : lo;
      v[q] = keep + __shfl_xor(send, step, kWave);
    }
  }
#pragma unroll
  for (int step = LG >> (K + 1); step >= 1; step >>= 1) v[0] += __shfl_xor(v[0], step, kWave);
}

template <int NV, int LG>
__device__ inline int group_value_slot(int l) {
  constexpr int KL = ilog2(LG), KV = ilog2(NV), K = KL < KV ? KL : KV;
  int q = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) q |= ((l >> (KL - 1 - k)) & 1) << (K - 1 - k);
  return q;
}

// FP32-pair / FP64-accumulate cluster kernel.  G i-clusters per wave, LG = 64/G lanes per i-cluster; the
// lanes of a group stream the flat j sequence of their i-cluster's list (entry e, member b) LG at a time
// and evaluate each j against the M i-particles.  Loads of the next j are issued before the pair math
// of the current one.  Padding j-members (and lanes past the end of the list) sit at +3e30 and padding
// i-members are moved to -3e30, so their r^2 overflows to +inf and the bit-pattern test
// (bits(r^2) - 1) < cut rejects them and the i == j self pair (r^2 = +0 wraps to the largest value).
// INTEG: the velocity-Verlet update in the epilogue, as in k_force_cluster_jrows (kick of this step, drift of
// the next, FP64 in k_kick_drift's operation order); the next FP32 cluster-relative positions go to
// ia.pos32_next and the next FP64 positions to ia.posd_next, because other workgroups still read this
// step's.
// The cutoff decision is the FP64 one, as in k_force_cluster_jrows: the FP32 r^2 is compared with two
// thresholds around rc^2 (cut_lo, cut_hi), and a wave with any pair in between decides those pairs again
// from the FP64 positions of this step (bd.posd) with the entry's periodic shift, exactly as
// k_force_cluster_fp64 does (r^2 <= rc^2 with r^2 > 0).  The band is four times a bound on the FP32 error
// of r^2, taken per wave from the largest cluster-relative offset since the last build (bd.ext) plus the
// skin/2 a particle may move before the next rebuild: the separation x_q - (p_j + off) rounds each of
// its four operands once and its two sums once, every operand below 2E + r for members within E of
// their cluster's centre, so |error| <= 2^-24 (5E + 3r) per component.
struct MixedBand {
  const double4 *posd;             // FP64 positions of this step, per slot
  const unsigned long long *ext;   // kAccSlots lines: the largest |pos32 component| (float bits)
  double L[3];                     // box of a single-box context
  const TowerGeom *btg;            // ensemble: per-box geometry (its L), else null
  const int32_t *cluster_box;      // ensemble: box of each cluster
  double rc2;
  float rc, half_skin;
};

// N3 (Newton-3 lists, single boxes): every listed pair is evaluated once and its force goes to both
// particles (LJFunctor.h:316-318): each lane sums its j's reaction over the M i-members of the entry and
// adds it to the j slot atomically; the i-members' sums are added atomically at the end (the force arrays
// were cleared before the launch).  On the cluster's own entry (shift id 13) only member pairs b > q count,
// as SoAFunctorSingle's triangle (LJFunctor.h:204-364).  Pair hits, energy and virial count twice, so that
// the common halving of the full-list sums (LJFunctor::endTraversal) gives the pair values.
// XR (member rows): the lanes stream the i-cluster's member row (xcount[ic] elements {j slot, entry} at
// xrow[ic * xstride]: the listed j-cluster members within rc + skin of one of its members) instead of
// every member of every listed entry; the element gives the j slot and its entry directly, so their two
// loads are issued together one iteration ahead.
template <int M, int G, bool MIX, bool GLOB, bool INTEG, bool N3, bool XR>
__global__ __launch_bounds__(kFThreads, (XR && !MIX && !GLOB) ? 4 : 1) void k_force_cluster_mixed(
    int64_t nclusters, Segs sg, const float4 *__restrict__ pos32, int32_t row, const int32_t *__restrict__ lcount,
    const int4 *__restrict__ list, const int32_t *__restrict__ xcount, const int2 *__restrict__ xrow,
    int32_t xstride, float rc2, float eps24_1, float sig2_1, float shift6_1, int32_t ntypes,
    const double *__restrict__ tab64, double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
    double *__restrict__ pe, double *__restrict__ pv, int64_t *__restrict__ ph, unsigned long long *__restrict__ acc,
    SkipIf sk, double *__restrict__ clg, IntegArgs ia, MixedBand bd, Owned own, WorkQueue wq) {
  static_assert(!(N3 && INTEG), "the fused update needs every pair's force first");
  static_assert(!(N3 && XR), "Newton-3 lists stream whole entries");
  if (skip_launch(sk)) return;
  constexpr int LG = kWave / G;
  constexpr int NVAL = 3 * M + 1 + (GLOB ? 2 : 0);
  constexpr int NV = pow2_at_least(NVAL);
  constexpr float kW = N3 ? 2.f : 1.f;  // weight of a hit in the pair count, energy and virial sums
  __shared__ float stab[MIX ? 3 * kMaxTypes * kMaxTypes : 1];
  if constexpr (MIX) {
    for (int k = threadIdx.x; k < 3 * ntypes * ntypes; k += kFThreads) stab[k] = (float)tab64[k];
    __syncthreads();
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int grp = lane / LG, l = lane % LG;
  // the FP64 band around rc^2 (uniform over the launch)
  uint32_t cut_lo, cut_hi;
  {
    float E = 0.f;
#pragma unroll
    for (int k = 0; k < kAccSlots; ++k) E = fmaxf(E, __uint_as_float((uint32_t)bd.ext[k * kAccStride]));
    E += bd.half_skin;
    const float u = 5.9604645e-8f;  // 2^-24
    const float rel = 4.f * (3.4641016f * u * (5.f * E + 3.f * bd.rc) / bd.rc + 3.f * u);
    const uint32_t w = (uint32_t)fminf(ceilf(rel * 8388608.f), 1.0e6f);
    const uint32_t rb = __float_as_uint(rc2);
    cut_lo = rb - w;
    cut_hi = rb + w;
  }
  double e_out = 0., v_out = 0.;
  int64_t h_out = 0;
  // Tiles of this wavefront (G i-clusters each): its place in the grid, or (work queue) chunks of
  // consecutive tiles taken from the regions' counters, its XCD's region first (wave-uniform values).
  const int wave_tile = (int)(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave);
  const int nchunks = wq.ctr ? (wq.ntiles + wq.chunk - 1) / wq.chunk : 0;
  int region = (int)(blockIdx.x % kQueueRegions), tries = 0;
  int tile = wq.ctr ? -1 : wave_tile, tile_end = tile + 1;
  for (;;) {
    if (wq.ctr) {
      ++tile;
      while (tile >= tile_end && tries < kQueueRegions) {
        const int r0 = (int)((int64_t)nchunks * region / kQueueRegions);
        const int r1 = (int)((int64_t)nchunks * (region + 1) / kQueueRegions);
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(wq.ctr + region, 1u);
        t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
        if ((int64_t)r0 + t < r1) {
          tile = (r0 + (int)t) * wq.chunk;
          tile_end = min(tile + wq.chunk, wq.ntiles);
        } else {
          region = (region + 1) % kQueueRegions;
          ++tries;
        }
      }
      if (tile >= tile_end) break;
    } else if (tile >= tile_end) {
      break;
    }
    const int item = tile * G + grp;
    int ic = item < nclusters ? (int)seg_index(sg, item) : -1;
    if constexpr (G == 1) ic = __builtin_amdgcn_readfirstlane(ic);
    if (!wq.ctr) ++tile;
  if (ic >= 0) {
    float xi[M], yi[M], zi[M];
    int ti[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const float4 p = pos32[(int64_t)ic * M + q];
      const bool pad = p.x == kPadPos32;
      xi[q] = pad ? -kPadPos32 : p.x;
      yi[q] = pad ? -kPadPos32 : p.y;
      zi[q] = pad ? -kPadPos32 : p.z;
      ti[q] = __float_as_int(p.w);
    }
    uint32_t real = 0;  // bit q set for the i-cluster's real (non-padding) members
#pragma unroll
    for (int q = 0; q < M; ++q) real |= (xi[q] == -kPadPos32 ? 0u : 1u) << q;
    float ax[M], ay[M], az[M], ae[M], av[M];
#pragma unroll
    for (int q = 0; q < M; ++q) { ax[q] = ay[q] = az[q] = ae[q] = av[q] = 0.f; }
    int hits = 0;
    const int64_t beg = (int64_t)ic * row;
    const int2 *xr = XR ? xrow + (int64_t)ic * xstride : nullptr;
    // >= M without member rows: every list holds its own cluster (a member row at least its real members)
    const int nj = XR ? xcount[ic] : lcount[ic] * M;
    // prefetch the first j of this lane (an empty list skips the loop and never loads), and with member
    // rows the element of the next one
    const int kk = l < nj ? l : nj - 1;
    int4 e = make_int4(0, 0, 0, 0);
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    int2 xq = make_int2(0, 0);
    if (nj > 0) {
      if constexpr (XR) {
        const int2 x0 = xr[kk];
        e = list[beg + x0.y];
        p = pos32[x0.x];
        xq = xr[l + LG < nj ? l + LG : nj - 1];
      } else {
        e = list[beg + (uint32_t)kk / M];
        p = pos32[(int64_t)(e.x & kClusterMask) * M + (kk & (M - 1))];
      }
    }
    for (int base = 0; base < nj; base += LG) {
      const bool valid = base + l < nj;
      const float xj = valid ? p.x + __int_as_float(e.y) : kPadPos32;
      const float yj = p.y + __int_as_float(e.z);
      const float zj = p.z + __int_as_float(e.w);
      const int tj = __float_as_int(p.w);
      // Newton-3: this lane's j slot, and whether its entry is the cluster's own (member pairs b > q only)
      int64_t sj_n3 = 0;
      bool own_n3 = false;
      const int bm = (base + l) & (M - 1);
      if constexpr (N3) {
        const int jc = e.x & kClusterMask;
        sj_n3 = (int64_t)jc * M + bm;
        own_n3 = jc == ic && ((uint32_t)e.x >> kShiftBits) == 13u;
      }
      float fjx = 0.f, fjy = 0.f, fjz = 0.f;  // Newton-3: the reaction on j, summed over the i-members
      // issue the next loads before the pair math (with member rows, also the element of the one after)
      if constexpr (XR) {
        e = list[beg + xq.y];
        p = pos32[xq.x];
        const int k2 = base + 2 * LG + l;
        xq = xr[k2 < nj ? k2 : nj - 1];
      } else {
        const int kn = base + LG + l;
        const int kc = kn < nj ? kn : nj - 1;
        e = list[beg + (uint32_t)kc / M];
        p = pos32[(int64_t)(e.x & kClusterMask) * M + (kc & (M - 1))];
      }
      auto pair_term = [&](int q, float dx, float dy, float dz, float r2, bool m) {
        float eps24 = eps24_1, sig2 = sig2_1, shift6 = shift6_1;
        if constexpr (MIX) {
          const int t = 3 * (ti[q] * ntypes + tj);
          eps24 = stab[t]; sig2 = stab[t + 1]; shift6 = stab[t + 2];
        }
        const float inv = __builtin_amdgcn_rcpf(r2);
        const float lj2 = sig2 * inv;
        const float lj6 = lj2 * lj2 * lj2;
        // eps24 (lj12 + lj12m6) / r^2 = (lj6 / r^2) (2 eps24 lj6 - eps24)
        const float fv = (lj6 * inv) * __builtin_fmaf(eps24 + eps24, lj6, -eps24);
        const float fac = m ? fv : 0.f;
        ax[q] = __builtin_fmaf(dx, fac, ax[q]);
        ay[q] = __builtin_fmaf(dy, fac, ay[q]);
        az[q] = __builtin_fmaf(dz, fac, az[q]);
        if constexpr (N3) {
          fjx = __builtin_fmaf(dx, fac, fjx);
          fjy = __builtin_fmaf(dy, fac, fjy);
          fjz = __builtin_fmaf(dz, fac, fjz);
        }
        hits += m ? (N3 ? 2 : 1) : 0;
        if constexpr (GLOB) {
          ae[q] += m ? kW * __builtin_fmaf(eps24, lj6 * lj6 - lj6, shift6) : 0.f;
          av[q] = __builtin_fmaf(m ? kW * r2 : 0.f, fac, av[q]);  // r2 is +inf for padding
        }
      };
      // pairs in the band around rc^2 are left out here; anyb collects the lanes that hold one
      unsigned long long anyb = 0ull;
#pragma unroll
      for (int q = 0; q < M; ++q) {
        const float dx = xi[q] - xj, dy = yi[q] - yj, dz = zi[q] - zj;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const uint32_t ub = __float_as_uint(r2) - 1u;
        const bool tri = !N3 || !own_n3 || bm > q;
        const bool m = ub < cut_lo && tri;
        anyb |= __builtin_amdgcn_ballot_w64(ub < cut_hi && tri) & ~__builtin_amdgcn_ballot_w64(m);
        pair_term(q, dx, dy, dz, r2, m);
      }
      if (anyb != 0ull) {
        // rare: find the band pairs again (same operands, same r^2) and decide them from the FP64
        // positions.  The j values pass through an empty asm and this lane's entry is loaded again, so
        // that the loop keeps the registers it needs without the band (as in k_force_cluster_jrows).
        float jx = xj, jy = yj, jz = zj;
        asm volatile("" : "+v"(jx), "+v"(jy), "+v"(jz));
        const int kc = valid ? base + l : nj - 1;
        int4 ec;
        int64_t sj;
        if constexpr (XR) {
          const int2 xc = xr[kc];
          ec = list[beg + xc.y];
          sj = xc.x;
        } else {
          ec = list[beg + (uint32_t)kc / M];
          sj = (int64_t)(ec.x & kClusterMask) * M + (kc & (M - 1));
        }
        const int sid = (int)((uint32_t)ec.x >> kShiftBits);
        // the box lengths are read here, not held through the loop (six registers the common path needs)
        double Lx = bd.L[0], Ly = bd.L[1], Lz = bd.L[2];
        if (bd.btg) {
          const TowerGeom &g = bd.btg[bd.cluster_box[ic]];
          Lx = g.L[0]; Ly = g.L[1]; Lz = g.L[2];
        }
        const double4 pj = bd.posd[sj];
        double xjd = pj.x + shift_component(sid, 0) * Lx;
        double yjd = pj.y + shift_component(sid, 1) * Ly;
        double zjd = pj.z + shift_component(sid, 2) * Lz;
        uint32_t inside = 0u;
#pragma unroll
        for (int q = 0; q < M; ++q) {
          const float dx = xi[q] - jx, dy = yi[q] - jy, dz = zi[q] - jz;
          const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
          const uint32_t ub = __float_as_uint(r2) - 1u;
          const bool inband = ub >= cut_lo && ub < cut_hi && (!N3 || !own_n3 || bm > q);
          if (__ballot(inband) == 0ull) continue;
          int32_t qq = q;
          asm volatile("" : "+v"(qq));
          const double4 pi = bd.posd[(int64_t)ic * M + qq];
          const double drx = pi.x - xjd, dry = pi.y - yjd, drz = pi.z - zjd;
          const double dr2 = drx * drx + dry * dry + drz * drz;
          if (inband && dr2 <= bd.rc2 && dr2 > 0.) inside |= 1u << q;
          asm volatile("" : "+v"(jx), "+v"(jy), "+v"(jz), "+v"(xjd), "+v"(yjd), "+v"(zjd), "+v"(inside));
        }
#pragma unroll
        for (int q = 0; q < M; ++q) {
          const bool in = (inside >> q) & 1u;
          if (__ballot(in) != 0ull) {
            const float dx = xi[q] - jx, dy = yi[q] - jy, dz = zi[q] - jz;
            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
            pair_term(q, dx, dy, dz, r2, in);
          }
          asm volatile("" : "+v"(jx), "+v"(jy), "+v"(jz), "+v"(ax[q]), "+v"(ay[q]), "+v"(az[q]));
        }
      }
      if constexpr (N3) {
        // the reaction on this lane's j (padding and lanes past the row have none)
        if (valid && (fjx != 0.f || fjy != 0.f || fjz != 0.f)) {
          unsafeAtomicAdd(&fx[sj_n3], -(double)fjx);
          unsafeAtomicAdd(&fy[sj_n3], -(double)fjy);
          unsafeAtomicAdd(&fz[sj_n3], -(double)fjz);
        }
      }
    }
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.;
#pragma unroll
    for (int q = 0; q < M; ++q) {
      v[q] = (double)ax[q];
      v[M + q] = (double)ay[q];
      v[2 * M + q] = (double)az[q];
    }
    v[3 * M] = (double)hits;
    if constexpr (GLOB) {
      double se = 0., sv = 0.;
#pragma unroll
      for (int q = 0; q < M; ++q)
        if (is_owned(own, (int64_t)ic * M + q)) { se += (double)ae[q]; sv += (double)av[q]; }
      v[3 * M + 1] = se;
      v[3 * M + 2] = sv;
    }
    group_reduce_scatter<NV, LG>(v, l);
    constexpr int KL = ilog2(LG), KV = ilog2(NV), K = KL < KV ? KL : KV;
    constexpr int R = NV >> K;
    const int q0 = group_value_slot<NV, LG>(l) * R;
    if ((l & ((1 << (KL - K)) - 1)) == 0) {
      const int64_t s0 = (int64_t)ic * M;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = q0 + r;
        const double val = v[r];
        // padding members are never stored; Newton-3 adds to what other workgroups' reactions left
        if (q < 3 * M) {
          if (!INTEG && ((real >> (q % M)) & 1u)) {
            double *dst = (q < M ? fx : q < 2 * M ? fy : fz) + s0 + q % M;
            if constexpr (N3) unsafeAtomicAdd(dst, val);
            else *dst = val;
          }
        } else if (q == 3 * M) {
          h_out += (int64_t)val;
          if (GLOB && clg) clg[3 * (int64_t)ic + 2] = val;  // per-cluster sums for per-box globals
        } else if (GLOB && q == 3 * M + 1) {
          e_out += val;
          if (clg) clg[3 * (int64_t)ic] = val;
        } else if (GLOB && q == 3 * M + 2) {
          v_out += val;
          if (clg) clg[3 * (int64_t)ic + 1] = val;
        }
      }
    }
    if constexpr (INTEG) {
      // the group's forces go through LDS so that each lane then advances one whole member
      __shared__ double sF[kWavesPerBlock * G * 3 * M];
      double *fsh = sF + ((threadIdx.x / kWave) * G + grp) * (3 * M);
      if ((l & ((1 << (KL - K)) - 1)) == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int q = q0 + r;
          if (q < 3 * M) fsh[q] = v[r];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const double4 cc = ia.c0[ic];
      for (int m = l; m < M; m += LG) {
#pragma clang fp contract(off)
        if (!((real >> m) & 1u)) continue;
        const int64_t i = (int64_t)ic * M + m;
        const double Fx = fsh[m], Fy = fsh[M + m], Fz = fsh[2 * M + m];
        const int t = __float_as_int(pos32[i].w);
        fx[i] = Fx;
        fy[i] = Fy;
        fz[i] = Fz;
        double vx = ia.vx[i], vy = ia.vy[i], vz = ia.vz[i];
        if (ia.kick) {
          const double kf = ia.kick_fac[t];
          vx += (Fx + ia.ofx[i]) * kf;
          vy += (Fy + ia.ofy[i]) * kf;
          vz += (Fz + ia.ofz[i]) * kf;
          ia.vx[i] = vx; ia.vy[i] = vy; ia.vz[i] = vz;
        }
        if (ia.drift) {
          const double ff = ia.drift_fac[t];
          const double dx = vx * ia.dt + Fx * ff;
          const double dy = vy * ia.dt + Fy * ff;
          const double dz = vz * ia.dt + Fz * ff;
          ia.ofx[i] = Fx; ia.ofy[i] = Fy; ia.ofz[i] = Fz;
          const double x = ia.x[i] + dx, y = ia.y[i] + dy, z = ia.z[i] + dz;
          ia.x[i] = x; ia.y[i] = y; ia.z[i] = z;
          const double ex = x - ia.rx[i], ey = y - ia.ry[i], ez = z - ia.rz[i];
          if (ex * ex + ey * ey + ez * ez >= ia.half_skin2) {
            ia.flag[0] = ia.stamp;
            ia.hflag[0] = ia.stamp;
            if (ia.box_flag) ia.box_flag[ia.cluster_box[i / M]] = 1;
          }
          ia.pos32_next[i] = make_float4((float)(x - cc.x), (float)(y - cc.y), (float)(z - cc.z), __int_as_float(t));
          ia.posd_next[i] = make_double4(x, y, z, 0.);
        }
      }
      // the next tile's sums go to the same LDS slots
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  }
  if (wq.ctr && lane == 0) {
    // the last wavefront done clears the counters for the next launch (every other one has taken its
    // last tile and failed on every region by then)
    __threadfence();
    if (atomicAdd(wq.ctr + kQueueRegions, 1u) == (uint32_t)wq.nwaves - 1u) {
      for (int r = 0; r <= kQueueRegions; ++r) atomicExch(wq.ctr + r, 0u);
      __threadfence();
    }
  }
  block_partials(e_out, v_out, h_out, pe, pv, ph, GLOB, acc);
}

// i-clusters against j-particle rows on fixed-point positions.  G i-clusters per wave, LG = 64/G lanes per
// i-cluster; the lanes of a group stream the j slots of their row (the union of the members' neighbour
// lists, so no j is evaluated that is farther than rc + skin from every i-member) and evaluate each j
// against the M i-members held in registers.
//  * FAST (cubic boxes whose rows passed the frame check of the list build): positions enter as wrapping
//    int32 differences to the cluster's first member, converted to FP32 once per particle, and
//    x_q - x_j is one FP32 subtraction.  Everything runs in fixed-point units (r^2, sigma^2, rc^2 scaled
//    by (2^32/L)^2) and the force sums are rescaled once; energy and virial are scale-free.  Padding
//    i-members sit at -3e30 and lanes past the end of the row at +3e30, so their r^2 overflows to +inf
//    and fails the bit-pattern cutoff test.
//  * otherwise: every pair separation is the exact minimum-image int32 difference x_q - x_j, scaled per
//    dimension; padding members and lanes past the end of the row are masked.
//  * INTEG: the epilogue also advances the i-members by velocity Verlet exactly as k_kick_drift does
//    (kick of this step, drift of the next, same FP64 operations in the same order), writes their
//    fixed-point positions for the next step to the second buffer (other workgroups still read this
//    step's), and raises the displacement flag.
//  * The cutoff decision is the FP64 one.  The FP32 r^2 is compared with two thresholds around rc^2
//    (cut_lo, cut_hi, bit patterns of positive floats, wider apart than the FP32 error of r^2): below
//    cut_lo the pair is inside, from cut_hi on outside, and a wave with any pair in between decides those
//    pairs again in FP64 from the exact fixed-point separation (r^2 <= rc2d).  An FP32 decision that
//    differs from the FP64 one leaves the force at rc (~0.04 epsilon/sigma) on a particle, and the two
//    rows of a pair round r^2 differently in the FAST frame: at 10^6 particles a dozen such pairs would
//    dominate the force error.  The band holds ~3e-5 of the candidate pairs (~1% of wave iterations).
template <int M, int G, bool FAST, bool MIX, bool GLOB, bool INTEG>
__global__ __launch_bounds__(kFThreads) void k_force_cluster_jrows(
    int64_t nclusters, const int4 *__restrict__ posfx, const double4 *__restrict__ c0, int32_t row,
    const int32_t *__restrict__ lcount, const int32_t *__restrict__ jlist, FixGeom f, uint32_t cut_lo,
    uint32_t cut_hi, double rc2d, float eps24_1, float sig2_1, float shift6_1, int32_t ntypes,
    const double *__restrict__ tab64, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz, double *__restrict__ pe, double *__restrict__ pv,
    int64_t *__restrict__ ph, unsigned long long *__restrict__ acc, SkipIf sk, IntegArgs ia, Owned own) {
  if (skip_launch(sk)) return;
  constexpr int LG = kWave / G;
  constexpr int NVAL = 3 * M + (GLOB ? 2 : 0);
  constexpr int NV = pow2_at_least(NVAL);
  __shared__ float stab[MIX ? 3 * kMaxTypes * kMaxTypes : 1];
  if constexpr (MIX) {
    const float s2u = FAST ? (float)(f.inv_scale[0] * f.inv_scale[0]) : 1.f;
    for (int k = threadIdx.x; k < 3 * ntypes * ntypes; k += kFThreads)
      stab[k] = (float)tab64[k] * (k % 3 == 1 ? s2u : 1.f);
    __syncthreads();
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int grp = lane / LG, l = lane % LG;
  int ic = (int)(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave) * G + grp;
  if constexpr (G == 1) ic = __builtin_amdgcn_readfirstlane(ic);
  const bool live = ic < nclusters;
  const int64_t icl = live ? ic : nclusters - 1;  // dead groups (last wave only) read a real cluster, store nothing
  const int nreal = (int)c0[icl].w;
  const int4 ref = posfx[icl * M];  // the first member is always real
  float xi[M], yi[M], zi[M];     // FAST: FP32 offsets from the first member (fixed-point units)
  uint32_t ux[M], uy[M], uz[M];  // exact variant: fixed-point positions
  int ti[M];
#pragma unroll
  for (int q = 0; q < M; ++q) {
    const int4 p = posfx[icl * M + q];
    const bool pad = q >= nreal;
    ux[q] = (uint32_t)p.x;
    uy[q] = (uint32_t)p.y;
    uz[q] = (uint32_t)p.z;
    xi[q] = pad ? -kPadPos32 : (float)(int32_t)((uint32_t)p.x - (uint32_t)ref.x);
    yi[q] = pad ? -kPadPos32 : (float)(int32_t)((uint32_t)p.y - (uint32_t)ref.y);
    zi[q] = pad ? -kPadPos32 : (float)(int32_t)((uint32_t)p.z - (uint32_t)ref.z);
    ti[q] = pad ? 0 : p.w;
  }
  float ax[M], ay[M], az[M], ae[M], av[M];
#pragma unroll
  for (int q = 0; q < M; ++q) { ax[q] = ay[q] = az[q] = ae[q] = av[q] = 0.f; }
  // the pair count is kept per wavefront on the scalar unit: popcount of each ballot of the cutoff mask
  uint32_t hits = 0;
  const int32_t *rowp = jlist + icl * row;
  const int nj = live ? lcount[icl] : 0;  // >= nreal for live groups: every row holds the cluster's own members
  // the loop runs to the wave's longest row so that it stays uniform (the ballot count needs that)
  int njmax = nj;
  if constexpr (G > 1) {
#pragma unroll
    for (int off = 32; off >= LG; off >>= 1) {
      const int o = __shfl_xor(njmax, off, kWave);
      njmax = o > njmax ? o : njmax;
    }
  }
  const int jlast = nj > 0 ? nj - 1 : 0;
  // every cluster's row holds its own members, so entry jlast exists for every group (dead groups read the
  // row of a real cluster): the loads need no guard
  auto load_j = [&](int k) { return posfx[rowp[k < nj ? k : jlast]]; };
  // one j per lane of the group against the M members: the j at base + l
  auto j_step = [&](int base, const int4 &pj) {
    const bool valid = base + l < nj;
    const int tj = pj.w < 0 ? 0 : pj.w;
    float xj = 0.f, yj = 0.f, zj = 0.f;
    if constexpr (FAST) {
      xj = valid ? (float)(int32_t)((uint32_t)pj.x - (uint32_t)ref.x) : kPadPos32;
      yj = (float)(int32_t)((uint32_t)pj.y - (uint32_t)ref.y);
      zj = (float)(int32_t)((uint32_t)pj.z - (uint32_t)ref.z);
    }
    auto separation = [&](int q, float jx, float jy, float jz, const int4 &pp, float &dx, float &dy, float &dz) {
      if constexpr (FAST) {
        dx = xi[q] - jx; dy = yi[q] - jy; dz = zi[q] - jz;
      } else {
        dx = (float)(int32_t)(ux[q] - (uint32_t)pp.x) * f.scale32[0];
        dy = (float)(int32_t)(uy[q] - (uint32_t)pp.y) * f.scale32[1];
        dz = (float)(int32_t)(uz[q] - (uint32_t)pp.z) * f.scale32[2];
      }
    };
    auto pair_term = [&](int q, float dx, float dy, float dz, float r2, bool m) {
      if constexpr (!FAST) m = m && valid && q < nreal;
      float eps24 = eps24_1, sig2 = sig2_1, shift6 = shift6_1;
      if constexpr (MIX) {
        const int t = 3 * (ti[q] * ntypes + tj);
        eps24 = stab[t]; sig2 = stab[t + 1]; shift6 = stab[t + 2];
      }
      const float inv = __builtin_amdgcn_rcpf(r2);
      const float lj2 = sig2 * inv;
      const float lj6 = lj2 * lj2 * lj2;
      const float fv = (lj6 * inv) * __builtin_fmaf(eps24 + eps24, lj6, -eps24);
      const float fac = m ? fv : 0.f;
      ax[q] = __builtin_fmaf(dx, fac, ax[q]);
      ay[q] = __builtin_fmaf(dy, fac, ay[q]);
      az[q] = __builtin_fmaf(dz, fac, az[q]);
      hits += (uint32_t)__popcll(__ballot(m));
      if constexpr (GLOB) {
        ae[q] += m ? __builtin_fmaf(eps24, lj6 * lj6 - lj6, shift6) : 0.f;
        av[q] = __builtin_fmaf(m ? r2 : 0.f, fac, av[q]);
      }
    };
    // pairs in the band around rc^2 are left out here; anyb collects the lanes that hold one
    unsigned long long anyb = 0ull;
#pragma unroll
    for (int q = 0; q < M; ++q) {
      float dx, dy, dz;
      separation(q, xj, yj, zj, pj, dx, dy, dz);
      const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
      const uint32_t ub = __float_as_uint(r2) - 1u;  // r2 = +0 wraps: the pair of a member with itself
      const bool m = ub < cut_lo;
      // lane masks of the two compares, combined on the scalar unit
      anyb |= __builtin_amdgcn_ballot_w64(ub < cut_hi) & ~__builtin_amdgcn_ballot_w64(m);
      pair_term(q, dx, dy, dz, r2, m);
    }
    if (anyb != 0ull) {
      // rare: find the band pairs again (same operands, same r^2), decide them in FP64 from the exact
      // fixed-point separation and add those inside.  Nothing is kept from the loop above for this: the
      // j values pass through an empty asm and the positions are loaded again (the member's index also
      // through the asm, so that the load is not hoisted out of the j loop), which keeps the loop at
      // the registers it needs without the band.
      float jx = xj, jy = yj, jz = zj;
      asm volatile("" : "+v"(jx), "+v"(jy), "+v"(jz));
      int4 pp = make_int4(0, 0, 0, 0);
      if (nj > 0) pp = posfx[rowp[valid ? base + l : jlast]];
      // first the decisions, one bit per member, then the pair terms: the FP64 temporaries are dead
      // before the pair term needs its own
      uint32_t inside = 0u;
#pragma unroll
      for (int q = 0; q < M; ++q) {
        float dx, dy, dz;
        separation(q, jx, jy, jz, pp, dx, dy, dz);
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const uint32_t ub = __float_as_uint(r2) - 1u;
        const bool inband = ub >= cut_lo && ub < cut_hi;
        if (__ballot(inband) == 0ull) continue;
        int32_t qq = q;
        asm volatile("" : "+v"(qq));
        const int4 pq = posfx[icl * M + qq];
        const double ex = (double)(int32_t)((uint32_t)pq.x - (uint32_t)pp.x) * f.scale[0];
        const double ey = (double)(int32_t)((uint32_t)pq.y - (uint32_t)pp.y) * f.scale[1];
        const double ez = (double)(int32_t)((uint32_t)pq.z - (uint32_t)pp.z) * f.scale[2];
        const double d2 = ex * ex + ey * ey + ez * ez;
        if (inband && d2 > 0.0 && d2 <= rc2d) inside |= 1u << q;
        // one member at a time: the next one's operands come out of this asm, after this one's result
        asm volatile("" : "+v"(jx), "+v"(jy), "+v"(jz), "+v"(pp.x), "+v"(pp.y), "+v"(pp.z), "+v"(inside));
      }
#pragma unroll
      for (int q = 0; q < M; ++q) {
        const bool in = (inside >> q) & 1u;
        if (__ballot(in) != 0ull) {
          float dx, dy, dz;
          separation(q, jx, jy, jz, pp, dx, dy, dz);
          const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
          pair_term(q, dx, dy, dz, r2, in);
        }
        asm volatile("" : "+v"(jx), "+v"(jy), "+v"(jz), "+v"(pp.x), "+v"(pp.y), "+v"(pp.z), "+v"(ax[q]),
                     "+v"(ay[q]), "+v"(az[q]));
      }
    }
  };
  // two j steps per trip, each one's positions loaded during the other's: the prefetched values stay in
  // their registers across the back edge (a one-step loop copies them there every trip)
  int4 pa = load_j(l);
  for (int base = 0; base < njmax; base += 2 * LG) {
    const int4 pb = load_j(base + LG + l);
    j_step(base, pa);
    if (base + LG >= njmax) break;
    pa = load_j(base + 2 * LG + l);
    j_step(base + LG, pb);
  }
  // the sums pass through an empty asm: otherwise the compiler folds them to 0 for lanes that skip the
  // loop and materialises those zeros before it, 2 NV registers live for the whole loop
#pragma unroll
  for (int q = 0; q < M; ++q) {
    asm volatile("" : "+v"(ax[q]), "+v"(ay[q]), "+v"(az[q]));
    if constexpr (GLOB) asm volatile("" : "+v"(ae[q]), "+v"(av[q]));
  }
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.;
#pragma unroll
  for (int q = 0; q < M; ++q) {
    v[q] = (double)ax[q];
    v[M + q] = (double)ay[q];
    v[2 * M + q] = (double)az[q];
  }
  if constexpr (GLOB) {
    double se = 0., sv = 0.;
#pragma unroll
    for (int q = 0; q < M; ++q)
      if (is_owned(own, icl * M + q)) { se += (double)ae[q]; sv += (double)av[q]; }
    v[3 * M] = se;
    v[3 * M + 1] = sv;
  }
  group_reduce_scatter<NV, LG>(v, l);
  constexpr int KL = ilog2(LG), KV = ilog2(NV), K = KL < KV ? KL : KV;
  constexpr int R = NV >> K;
  const int q0 = group_value_slot<NV, LG>(l) * R;
  // FAST: the sums are in fixed-point units; F = sum(d_u * fac_u) * (2^32 / L)
  const double fs = FAST ? f.inv_scale[0] : 1.0;
  double e_out = 0., v_out = 0.;
  if (live && (l & ((1 << (KL - K)) - 1)) == 0) {
    const int64_t s0 = icl * M;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = q0 + r;
      const double val = v[r];
      if (q < M) {
        if (!INTEG && q < nreal) fx[s0 + q] = val * fs;
      } else if (q < 2 * M) {
        if (!INTEG && q - M < nreal) fy[s0 + q - M] = val * fs;
      } else if (q < 3 * M) {
        if (!INTEG && q - 2 * M < nreal) fz[s0 + q - 2 * M] = val * fs;
      } else if (GLOB && q == 3 * M) {
        e_out = val;
      } else if (GLOB && q == 3 * M + 1) {
        v_out = val;
      }
    }
  }
  if constexpr (INTEG) {
    // the group's forces go through LDS so that each lane then advances one whole member: the state
    // loads and stores of a wave cover consecutive slots, as in k_kick_drift
    __shared__ double sF[kWavesPerBlock * G * 3 * M];
    double *fsh = sF + ((threadIdx.x / kWave) * G + grp) * (3 * M);
    if ((l & ((1 << (KL - K)) - 1)) == 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = q0 + r;
        if (q < 3 * M) fsh[q] = v[r] * fs;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (live) {
      for (int m = l; m < nreal; m += LG) {
#pragma clang fp contract(off)
        const int64_t i = icl * M + m;
        const double Fx = fsh[m], Fy = fsh[M + m], Fz = fsh[2 * M + m];
        const int4 pf = posfx[i];
        const int t = pf.w;
        fx[i] = Fx;
        fy[i] = Fy;
        fz[i] = Fz;
        double vx = ia.vx[i], vy = ia.vy[i], vz = ia.vz[i];
        if (ia.kick) {
          const double kf = ia.kick_fac[t];
          vx += (Fx + ia.ofx[i]) * kf;
          vy += (Fy + ia.ofy[i]) * kf;
          vz += (Fz + ia.ofz[i]) * kf;
          ia.vx[i] = vx; ia.vy[i] = vy; ia.vz[i] = vz;
        }
        if (ia.drift) {
          const double ff = ia.drift_fac[t];
          const double dx = vx * ia.dt + Fx * ff;
          const double dy = vy * ia.dt + Fy * ff;
          const double dz = vz * ia.dt + Fz * ff;
          ia.ofx[i] = Fx; ia.ofy[i] = Fy; ia.ofz[i] = Fz;
          const double x = ia.x[i] + dx, y = ia.y[i] + dy, z = ia.z[i] + dz;
          ia.x[i] = x; ia.y[i] = y; ia.z[i] = z;
          const double ex = x - ia.rx[i], ey = y - ia.ry[i], ez = z - ia.rz[i];
          if (ex * ex + ey * ey + ez * ez >= ia.half_skin2) {
            ia.flag[0] = ia.stamp;
            ia.hflag[0] = ia.stamp;
          }
          ia.posfx_next[i] = make_int4((int32_t)(uint32_t)(uint64_t)__double2ll_rn((x - f.min[0]) * f.inv_scale[0]),
                                       (int32_t)(uint32_t)(uint64_t)__double2ll_rn((y - f.min[1]) * f.inv_scale[1]),
                                       (int32_t)(uint32_t)(uint64_t)__double2ll_rn((z - f.min[2]) * f.inv_scale[2]), t);
        }
      }
    }
  }
  block_partials(e_out, v_out, lane == 0 ? (int64_t)hits : 0, pe, pv, ph, GLOB, acc);
}

// FP64 cluster kernel: same tiles (one i-cluster per wave), absolute FP64 positions with the periodic
// shift of the entry, pair term exactly as LJFunctor.h:288-320.  N3: Newton-3 lists, as in
// k_force_cluster_mixed (reaction on j added atomically, own entry's member pairs b > q only).
template <int M, bool MIX, bool GLOB, bool N3>
__global__ __launch_bounds__(kFThreads) void k_force_cluster_fp64(
    int64_t nclusters, const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const int32_t *__restrict__ type, const int32_t *__restrict__ orig, int32_t row,
    const int32_t *__restrict__ lcount, const int4 *__restrict__ list, ForceArgs a, const double *__restrict__ tab, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz, double *__restrict__ pe, double *__restrict__ pv,
    int64_t *__restrict__ ph, unsigned long long *__restrict__ acc, SkipIf sk, Owned own) {
  if (skip_launch(sk)) return;
  constexpr int NVAL = 3 * M + 1 + (GLOB ? 2 : 0);
  constexpr int NV = pow2_at_least(NVAL);
  const int lane = threadIdx.x & (kWave - 1);
  const int ic = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave));
  double e_out = 0., v_out = 0.;
  int64_t h_out = 0;
  if (ic < nclusters) {
    double xi[M], yi[M], zi[M];
    int ti[M];
    uint32_t real = 0;  // bit q set for real (non-padding) members
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const int64_t s = (int64_t)ic * M + q;
      xi[q] = x[s]; yi[q] = y[s]; zi[q] = z[s];
      ti[q] = type[s];
      real |= (orig[s] >= 0 ? 1u : 0u) << q;
    }
    double ax[M], ay[M], az[M], ae[M], av[M];
#pragma unroll
    for (int q = 0; q < M; ++q) { ax[q] = ay[q] = az[q] = ae[q] = av[q] = 0.; }
    int hits = 0;
    const int64_t beg = (int64_t)ic * row;
    const int nj = lcount[ic] * M;
    for (int base = 0; base < nj; base += kWave) {
      const int k = base + lane;
      const bool valid = k < nj;
      const int kk = valid ? k : nj - 1;
      const int4 e = list[beg + (uint32_t)kk / M];
      const int sid = (int)((uint32_t)e.x >> kShiftBits);
      const int64_t sj = (int64_t)(e.x & kClusterMask) * M + (kk & (M - 1));
      const double xj = x[sj] + shift_component(sid, 0) * a.L[0];
      const double yj = y[sj] + shift_component(sid, 1) * a.L[1];
      const double zj = z[sj] + shift_component(sid, 2) * a.L[2];
      const int tj = type[sj];
      const bool rj = valid && orig[sj] >= 0;
      const int bm = kk & (M - 1);
      const bool own = N3 && (e.x & kClusterMask) == ic && sid == 13;
      double fjx = 0., fjy = 0., fjz = 0.;
#pragma unroll
      for (int q = 0; q < M; ++q) {
        const double drx = xi[q] - xj, dry = yi[q] - yj, drz = zi[q] - zj;
        const double dr2 = drx * drx + dry * dry + drz * drz;
        double eps24 = a.eps24, sig2 = a.sig2, shift6 = a.shift6;
        if constexpr (MIX) {
          const int t = 3 * (ti[q] * a.ntypes + tj);
          eps24 = tab[t]; sig2 = tab[t + 1]; shift6 = tab[t + 2];
        }
        const bool m = rj && ((real >> q) & 1u) && dr2 <= a.rc2 && dr2 > 0. && (!own || bm > q);
        const double invdr2 = 1. / dr2;
        const double lj2 = sig2 * invdr2;
        const double lj6 = lj2 * lj2 * lj2;
        const double lj12 = lj6 * lj6;
        const double lj12m6 = lj12 - lj6;
        const double fac = m ? eps24 * (lj12 + lj12m6) * invdr2 : 0.;
        ax[q] += drx * fac;
        ay[q] += dry * fac;
        az[q] += drz * fac;
        if constexpr (N3) {
          fjx += drx * fac;
          fjy += dry * fac;
          fjz += drz * fac;
        }
        hits += m ? (N3 ? 2 : 1) : 0;
        if constexpr (GLOB) {
          const double w = N3 ? 2. : 1.;
          ae[q] += m ? w * (eps24 * lj12m6 + shift6) : 0.;
          av[q] += w * (drx * (drx * fac) + dry * (dry * fac) + drz * (drz * fac));
        }
      }
      if constexpr (N3) {
        if (rj && (fjx != 0. || fjy != 0. || fjz != 0.)) {
          unsafeAtomicAdd(&fx[sj], -fjx);
          unsafeAtomicAdd(&fy[sj], -fjy);
          unsafeAtomicAdd(&fz[sj], -fjz);
        }
      }
    }
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.;
#pragma unroll
    for (int q = 0; q < M; ++q) {
      v[q] = ax[q];
      v[M + q] = ay[q];
      v[2 * M + q] = az[q];
    }
    v[3 * M] = (double)hits;
    if constexpr (GLOB) {
      double se = 0., sv = 0.;
#pragma unroll
      for (int q = 0; q < M; ++q)
        if (is_owned(own, (int64_t)ic * M + q)) { se += ae[q]; sv += av[q]; }
      v[3 * M + 1] = se;
      v[3 * M + 2] = sv;
    }
    group_reduce_scatter<NV, kWave>(v, lane);
    const int q = group_value_slot<NV, kWave>(lane);
    constexpr int KV = ilog2(NV);
    if ((lane & ((1 << (6 - KV)) - 1)) == 0) {
      const int64_t s0 = (int64_t)ic * M;
      if (q < 3 * M) {
        if ((real >> (q % M)) & 1u) {
          double *dst = (q < M ? fx : q < 2 * M ? fy : fz) + s0 + q % M;
          if constexpr (N3) unsafeAtomicAdd(dst, v[0]);
          else *dst = v[0];
        }
      } else if (q == 3 * M) {
        h_out = (int64_t)v[0];
      } else if (GLOB && q == 3 * M + 1) {
        e_out = v[0];
      } else if (GLOB && q == 3 * M + 2) {
        v_out = v[0];
      }
    }
  }
  block_partials(e_out, v_out, h_out, pe, pv, ph, GLOB, acc);
}

static ForceArgs make_args(const Ctx &c) {
  ForceArgs a{};
  for (int d = 0; d < 3; ++d) a.L[d] = c.prm.box_max[d] - c.prm.box_min[d];
  a.rc2 = c.lj.rc2;
  a.ntypes = c.lj.ntypes;
  a.eps24 = c.lj.eps24[0];
  a.sig2 = c.lj.sig2[0];
  a.shift6 = c.lj.shift6[0];
  return a;
}

void lc_force_launch(hipStream_t st, int64_t n, const double *x, const double *y, const double *z,
                     const int32_t *type, const int32_t *cell_of, const int64_t *cell_start, const CellGeom &g,
                     double rc2, int32_t ntypes, const double *tab, double eps24, double sig2, double shift6,
                     double *fx, double *fy, double *fz, double *pe, double *pv, int64_t *ph, bool globals,
                     unsigned long long *acc, SkipIf skip, Owned own) {
  ForceArgs a{};
  for (int d = 0; d < 3; ++d) a.L[d] = g.L[d];
  a.rc2 = rc2;
  a.ntypes = ntypes;
  a.eps24 = eps24;
  a.sig2 = sig2;
  a.shift6 = shift6;
  const int nb = blocks_for(n, kFThreads);
  if (ntypes > 1)
    hipLaunchKernelGGL(k_force_lc_fp64<true>, dim3(nb), dim3(kFThreads), 0, st, n, x, y, z, type, cell_of,
                       cell_start, g, a, tab, fx, fy, fz, pe, pv, ph, (int32_t)globals, acc, skip, own);
  else
    hipLaunchKernelGGL(k_force_lc_fp64<false>, dim3(nb), dim3(kFThreads), 0, st, n, x, y, z, type, cell_of,
                       cell_start, g, a, tab, fx, fy, fz, pe, pv, ph, (int32_t)globals, acc, skip, own);
  LJG_HIP(hipGetLastError());
}

void lc_force(Ctx &c, bool globals) {
  const bool mix = c.lj.ntypes > 1;
  const double *tab = mix ? c.tab.p : nullptr;
  const int nb = blocks_for(c.n, kFThreads);
  c.part_e.ensure(nb);
  c.part_v.ensure(nb);
  c.part_hits.ensure(nb);
  c.part_items = nb;
  State &s = c.st[c.cur];
  if (c.timing) LJG_HIP(hipEventRecord(c.ev0, c.stream));
  lc_force_launch(c.stream, c.n, s.x.p, s.y.p, s.z.p, s.type.p, c.cell_of.p, c.cell_start.p, c.cg, c.lj.rc2,
                  c.lj.ntypes, tab, c.lj.eps24[0], c.lj.sig2[0], c.lj.shift6[0], s.fx.p, s.fy.p, s.fz.p, c.part_e.p,
                  c.part_v.p, c.part_hits.p, globals, c.hits_acc, c.skip, owned_of(c));
  if (c.timing) LJG_HIP(hipEventRecord(c.ev1, c.stream));
}

__global__ void k_zero_forces(int64_t n, double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
                              SkipIf sk) {
  if (skip_launch(sk)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    fx[i] = 0.;
    fy[i] = 0.;
    fz[i] = 0.;
  }
}

template <int M, int G>
static void launch_mixed(Ctx &c, bool mix, bool glob, const ForceArgs &a, const double *tab, const IntegArgs *integ,
                         const ClusterSubset *sub, bool n3) {
  constexpr int per_block = kWavesPerBlock * G;
  const Segs sg = sub ? sub->segs : Segs{nullptr, nullptr, 0};
  const int64_t items = sub ? sub->nitems : c.nclusters;
  if (items == 0) return;
  const int nb = blocks_for(items, per_block);
  // a subset launch writes its block partials after part_offset (reserved beforehand: launches of two
  // subsets may run at once)
  const int64_t off = sub ? sub->part_offset : 0;
  if (!sub) {
    c.part_e.ensure(nb);
    c.part_v.ensure(nb);
    c.part_hits.ensure(nb);
    c.part_items = nb;
  } else if (c.part_e.cap < (size_t)(off + nb) || c.part_v.cap < (size_t)(off + nb) ||
             c.part_hits.cap < (size_t)(off + nb)) {
    throw HipError{hipErrorInvalidValue, "force partials not reserved for a subset launch", __LINE__};
  }
  const float rc2 = (float)a.rc2;
  double *clg = nullptr;
  if (glob && c.nboxes > 0) {
    c.cl_glob.ensure(3 * (size_t)c.nclusters);
    clg = c.cl_glob.p;
  }
  const IntegArgs ia = integ ? *integ : IntegArgs{};
  MixedBand bd{};
  bd.posd = c.posd.p;
  bd.ext = c.pos32_ext.p;
  for (int d = 0; d < 3; ++d) bd.L[d] = a.L[d];
  if (c.nboxes > 0) {
    bd.btg = c.box_tg.p;
    bd.cluster_box = c.cluster_box.p;
  }
  bd.rc2 = a.rc2;
  bd.rc = (float)std::sqrt(a.rc2);
  bd.half_skin = (float)(0.5 * c.prm.skin);
  // member rows when the build made them (every cluster-entry build but Newton-3 and whole rows)
  const bool xr = c.xrows_valid && !n3;
  // Persistent launch (the step loop's full launches): a few wavefronts per SIMD take the tiles from the
  // work queue, four consecutive tiles at a time; its counters are cleared once here and then by every
  // launch's last wavefront.  Block partials then depend on which tiles a block took, so launches that
  // sum globals keep a wavefront per tile.
  WorkQueue wq{nullptr, 0, 0, 0};
  const int64_t tiles = blocks_for(items, G);
  const int persist_blocks = 256 * 4;  // 256 CUs, 4 workgroups (16 wavefronts) each
  int launch_blocks = nb;
  if (c.prm.persistent && integ && !sub && !glob && !n3 && nb > persist_blocks) {
    if (!c.queue_ready) {
      c.queue.ensure(kQueueRegions + 1);
      fill_i32(c, reinterpret_cast<int32_t *>(c.queue.p), kQueueRegions + 1, 0);
      c.queue_ready = true;
    }
    wq = WorkQueue{c.queue.p, (int32_t)tiles, 4, persist_blocks * kWavesPerBlock};
    launch_blocks = persist_blocks;
    c.part_items = launch_blocks;  // the block partials written
  }
#define LJG_LAUNCH1(MIXV, GLOBV, INTV, N3V, XRV)                                                                     \
  hipLaunchKernelGGL((k_force_cluster_mixed<M, G, MIXV, GLOBV, INTV, N3V, XRV>), dim3(launch_blocks), dim3(kFThreads), \
                     0, c.stream, items, sg, c.pos32.p, c.list_row, c.list_count.p, c.list.p, c.xcount.p, c.xrow.p,  \
                     c.xstride, rc2, (float)a.eps24, (float)a.sig2, (float)a.shift6, a.ntypes, tab,                  \
                     c.st[c.cur].fx.p, c.st[c.cur].fy.p, c.st[c.cur].fz.p, c.part_e.p + off, c.part_v.p + off,       \
                     c.part_hits.p + off, c.hits_acc, c.skip, clg, ia, bd, owned_of(c), wq)
#define LJG_LAUNCH(MIXV, GLOBV, INTV, N3V)                                                                         \
  do {                                                                                                             \
    if constexpr (N3V) {                                                                                           \
      LJG_LAUNCH1(MIXV, GLOBV, INTV, N3V, false);                                                                  \
    } else if (xr) {                                                                                               \
      LJG_LAUNCH1(MIXV, GLOBV, INTV, N3V, true);                                                                   \
    } else {                                                                                                       \
      LJG_LAUNCH1(MIXV, GLOBV, INTV, N3V, false);                                                                  \
    }                                                                                                              \
  } while (0)
  if (n3) {
    State &s = c.st[c.cur];
    hipLaunchKernelGGL(k_zero_forces, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots, s.fx.p,
                       s.fy.p, s.fz.p, c.skip);
    if (mix) {
      if (glob) LJG_LAUNCH(true, true, false, true); else LJG_LAUNCH(true, false, false, true);
    } else {
      if (glob) LJG_LAUNCH(false, true, false, true); else LJG_LAUNCH(false, false, false, true);
    }
  } else if (integ) {
    if (mix) LJG_LAUNCH(true, false, true, false); else LJG_LAUNCH(false, false, true, false);
  } else if (mix) {
    if (glob) LJG_LAUNCH(true, true, false, false); else LJG_LAUNCH(true, false, false, false);
  } else {
    if (glob) LJG_LAUNCH(false, true, false, false); else LJG_LAUNCH(false, false, false, false);
  }
#undef LJG_LAUNCH
#undef LJG_LAUNCH1
}

template <int M>
static void launch_fp64(Ctx &c, bool mix, bool glob, const ForceArgs &a, const double *tab, bool n3) {
  const int nb = blocks_for(c.nclusters, kWavesPerBlock);
  c.part_e.ensure(nb);
  c.part_v.ensure(nb);
  c.part_hits.ensure(nb);
  c.part_items = nb;
  State &s = c.st[c.cur];
#define LJG_LAUNCH(MIXV, GLOBV, N3V)                                                                              \
  hipLaunchKernelGGL((k_force_cluster_fp64<M, MIXV, GLOBV, N3V>), dim3(nb), dim3(kFThreads), 0, c.stream,        \
                     c.nclusters, s.x.p, s.y.p, s.z.p, s.type.p, s.orig.p, c.list_row, c.list_count.p, c.list.p, a, \
                     tab, s.fx.p, s.fy.p, s.fz.p, c.part_e.p, c.part_v.p, c.part_hits.p, c.hits_acc, c.skip,    \
                     owned_of(c))
  if (n3) {
    hipLaunchKernelGGL(k_zero_forces, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots, s.fx.p,
                       s.fy.p, s.fz.p, c.skip);
    if (mix) {
      if (glob) LJG_LAUNCH(true, true, true); else LJG_LAUNCH(true, false, true);
    } else {
      if (glob) LJG_LAUNCH(false, true, true); else LJG_LAUNCH(false, false, true);
    }
  } else if (mix) {
    if (glob) LJG_LAUNCH(true, true, false); else LJG_LAUNCH(true, false, false);
  } else {
    if (glob) LJG_LAUNCH(false, true, false); else LJG_LAUNCH(false, false, false);
  }
#undef LJG_LAUNCH
}

template <int M, int G>
static void launch_jrows(Ctx &c, bool mix, bool glob, const ForceArgs &a, const double *tab, const IntegArgs *integ) {
  constexpr int per_block = kWavesPerBlock * G;
  const int nb = blocks_for(c.nclusters, per_block);
  c.part_e.ensure(nb);
  c.part_v.ensure(nb);
  c.part_hits.ensure(nb);
  c.part_items = nb;
  const FixGeom &f = c.fg;
  const bool fast = f.cubic && c.jrows_frame_ok;
  const double u2 = fast ? f.inv_scale[0] * f.inv_scale[0] : 1.0;  // (2^32/L)^2 in the fast variant
  const float rc2 = (float)(a.rc2 * u2);
  const float sig2 = (float)(a.sig2 * u2);
  // band around rc^2 decided in FP64, four times the FP32 error of r^2 and at least 2^-18 relative.  The
  // exact variant's r^2 is off by ~9 roundings (conversion, scale, three squares and sums: ~5e-7).
  // FAST offsets (up to 2^31 units from the first member) round by at most 64 units each, the
  // separation by at most 192, so r^2 is off by at most 2 sqrt(3) 192 / rc_u relative plus a few
  // roundings.  The band is kept narrow because a wave iteration holds 64 M candidate pairs: at 2^-12
  // one iteration in ten had a pair in the band and took the FP64 pass.
  double rel = std::ldexp(1.0, -18);
  if (fast) {
    const double rc_u = std::sqrt(a.rc2 * u2);
    rel = std::max(rel, 4.0 * (2.0 * std::sqrt(3.0) * 192.0 / rc_u + std::ldexp(4.0, -24)));
  }
  uint32_t cut_lo, cut_hi;
  cutoff_band(rc2, rel, cut_lo, cut_hi);
  State &s = c.st[c.cur];
  const IntegArgs ia = integ ? *integ : IntegArgs{};
  // rows whose lengths the host has not read: the launch exits on the device if they overflowed
  SkipIf sk = c.skip;
  if (c.lists_pending) {
    sk.lst = c.list_stats.p;
    sk.cap1 = (uint32_t)(c.pend_cap / kJclSplit);
    sk.cap2 = (uint32_t)c.pend_jcap;
    sk.frame = fast ? 1 : 0;
    c.pend_fast = fast;
  }
#define LJG_LAUNCH(CUB, MIXV, GLOBV, INTV)                                                                          \
  hipLaunchKernelGGL((k_force_cluster_jrows<M, G, CUB, MIXV, GLOBV, INTV>), dim3(nb), dim3(kFThreads), 0, c.stream, \
                     c.nclusters, c.posfx.p, c.c0.p, c.list_row, c.jcount.p, c.jlist.p, f, cut_lo, cut_hi, a.rc2, \
                     (float)a.eps24, sig2, (float)a.shift6, a.ntypes, tab, s.fx.p, s.fy.p, s.fz.p, c.part_e.p,    \
                     c.part_v.p, c.part_hits.p, c.hits_acc, sk, ia, owned_of(c))
#define LJG_LAUNCH2(CUB)                                                                                 \
  if (integ) {                                                                                           \
    if (mix) LJG_LAUNCH(CUB, true, false, true); else LJG_LAUNCH(CUB, false, false, true);              \
  } else if (mix) {                                                                                      \
    if (glob) LJG_LAUNCH(CUB, true, true, false); else LJG_LAUNCH(CUB, true, false, false);             \
  } else {                                                                                               \
    if (glob) LJG_LAUNCH(CUB, false, true, false); else LJG_LAUNCH(CUB, false, false, false);           \
  }
  if (fast) {
    LJG_LAUNCH2(true)
  } else {
    LJG_LAUNCH2(false)
  }
#undef LJG_LAUNCH2
#undef LJG_LAUNCH
}

// i-clusters per wave for the FP32 kernel: the tuned default per cluster size, or the caller's choice.
int default_groups(int M) { return M >= 8 ? 1 : M == 4 ? 2 : M == 2 ? 4 : 16; }

int64_t cluster_force_blocks(const Ctx &c, int64_t nitems) {
  const int G = c.prm.tile_groups > 0 ? c.prm.tile_groups : default_groups(c.M);
  return blocks_for(nitems, kWavesPerBlock * G);
}

void cluster_force(Ctx &c, bool mixed, bool globals, const IntegArgs *integ, const ClusterSubset *sub, bool n3) {
  if (integ && (!mixed || globals)) throw HipError{hipErrorInvalidValue, "fused update needs the FP32 kernel", -2};
  if (sub && !mixed) throw HipError{hipErrorInvalidValue, "cluster subsets need the FP32 kernel", -2};
  if (n3 && (integ || sub || c.nboxes > 0 || !c.lists_n3))
    throw HipError{hipErrorInvalidValue, "Newton-3 runs on the Newton-3 lists of a single box, without fused update", -2};
  if (!n3 && c.lists_n3) throw HipError{hipErrorInvalidValue, "Newton-3 lists need a Newton-3 kernel", -2};
  const ForceArgs a = make_args(c);
  const bool mix = c.lj.ntypes > 1;
  const double *tab = mix ? c.tab.p : nullptr;
  const int G = c.prm.tile_groups > 0 ? c.prm.tile_groups : default_groups(c.M);
  if (c.timing) LJG_HIP(hipEventRecord(c.ev0, c.stream));
  if (!mixed) {
    switch (c.M) {
      case 1: launch_fp64<1>(c, mix, globals, a, tab, n3); break;
      case 2: launch_fp64<2>(c, mix, globals, a, tab, n3); break;
      case 4: launch_fp64<4>(c, mix, globals, a, tab, n3); break;
      case 8: launch_fp64<8>(c, mix, globals, a, tab, n3); break;
      default: throw HipError{hipErrorInvalidValue, "cluster size must be 1, 2, 4 or 8", __LINE__};
    }
  } else {
#define LJG_CASE(MV, GV) \
  if (c.M == MV && G == GV) { launch_mixed<MV, GV>(c, mix, globals, a, tab, integ, sub, n3); goto launched; }
    LJG_CASE(8, 1) LJG_CASE(8, 2) LJG_CASE(8, 4) LJG_CASE(8, 8)
    LJG_CASE(4, 1) LJG_CASE(4, 2) LJG_CASE(4, 4) LJG_CASE(4, 8)
    LJG_CASE(2, 1) LJG_CASE(2, 2) LJG_CASE(2, 4) LJG_CASE(2, 8) LJG_CASE(2, 16)
    LJG_CASE(1, 4) LJG_CASE(1, 8) LJG_CASE(1, 16) LJG_CASE(1, 32) LJG_CASE(1, 64)
#undef LJG_CASE
    throw HipError{hipErrorInvalidValue, "unsupported (cluster size, i-clusters per wave) combination", -2};
  launched:;
  }
  LJG_HIP(hipGetLastError());
  if (c.timing) LJG_HIP(hipEventRecord(c.ev1, c.stream));
}

// i-clusters per wave for the j-row kernel (rows are ~1.5x longer than a member's own list).
// Small builds (few i-clusters) are bound by the latency of one wave's row, not by throughput: fewer
// i-clusters per wave spread the row over more lanes (measured for M = 8 from 512 to 64,000 particles).
static int default_jrow_groups(int M, int64_t nclusters) {
  if (M >= 8) return nclusters <= 2048 ? 1 : nclusters <= 8192 ? 2 : 4;
  return M == 4 ? 8 : M == 2 ? 16 : 32;
}

void cluster_jrows_force(Ctx &c, bool globals, const IntegArgs *integ) {
  const ForceArgs a = make_args(c);
  const bool mix = c.lj.ntypes > 1;
  const double *tab = mix ? c.tab.p : nullptr;
  const int G = c.prm.tile_groups > 0 ? c.prm.tile_groups : default_jrow_groups(c.M, c.nclusters);
  if (c.timing) LJG_HIP(hipEventRecord(c.ev0, c.stream));
#define LJG_CASE(MV, GV) \
  if (c.M == MV && G == GV) { launch_jrows<MV, GV>(c, mix, globals && !integ, a, tab, integ); goto launched; }
  LJG_CASE(8, 1) LJG_CASE(8, 2) LJG_CASE(8, 4) LJG_CASE(8, 8)
  LJG_CASE(4, 2) LJG_CASE(4, 4) LJG_CASE(4, 8) LJG_CASE(4, 16)
  LJG_CASE(2, 8) LJG_CASE(2, 16)
  LJG_CASE(1, 16) LJG_CASE(1, 32)
#undef LJG_CASE
  throw HipError{hipErrorInvalidValue, "unsupported (cluster size, i-clusters per wave) combination", -2};
launched:
  LJG_HIP(hipGetLastError());
  if (c.timing) LJG_HIP(hipEventRecord(c.ev1, c.stream));
}

}  // namespace ljg
