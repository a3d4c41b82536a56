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
__device__ inline void block_partials(double e, double v, int64_t h, double *pe, double *pv, int64_t *ph,
                                      bool globals) {
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
  }
}

// ---- plain FP64 linked-cells kernel ------------------------------------------------------------------
template <bool MIX>
__global__ __launch_bounds__(kFThreads) void k_force_lc_fp64(
    int64_t n, const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const int32_t *__restrict__ type, const int32_t *__restrict__ cell_of, const int64_t *__restrict__ cell_start,
    CellGeom g, ForceArgs a, const double *__restrict__ tab, double *__restrict__ fx, double *__restrict__ fy,
    double *__restrict__ fz, double *__restrict__ pe, double *__restrict__ pv, int64_t *__restrict__ ph,
    int32_t globals) {
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
  }
  block_partials(ae, av, hits, pe, pv, ph, globals != 0);
}

// ---- cluster-pair tiles -------------------------------------------------------------------------------
// Wave reduce-scatter: v[NV] per lane (NV a power of two <= 64).  After the call lane l holds, in v[0],
// the wave total of value index vidx<NV>(l).
template <int NV>
__device__ inline void wave_reduce_scatter(double (&v)[NV], int lane) {
  constexpr int K = NV == 1 ? 0 : NV == 2 ? 1 : NV == 4 ? 2 : NV == 8 ? 3 : NV == 16 ? 4 : NV == 32 ? 5 : 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int step = 32 >> k;
    const int half = NV >> (k + 1);
    const bool upper = (lane & step) != 0;
#pragma unroll
    for (int q = 0; q < half; ++q) {
      const double send = upper ? v[q] : v[q + half];
      const double keep = upper ? v[q + half] : v[q];
      v[q] = keep + __shfl_xor(send, step, kWave);
    }
  }
#pragma unroll
  for (int step = 32 >> K; step >= 1; step >>= 1) v[0] += __shfl_xor(v[0], step, kWave);
}

template <int NV>
__device__ inline int vidx(int lane) {
  constexpr int K = NV == 1 ? 0 : NV == 2 ? 1 : NV == 4 ? 2 : NV == 8 ? 3 : NV == 16 ? 4 : NV == 32 ? 5 : 6;
  int q = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) q |= ((lane >> (5 - k)) & 1) << (K - 1 - k);
  return q;
}

constexpr int pow2_at_least(int x) { return x <= 1 ? 1 : 2 * pow2_at_least((x + 1) / 2); }

template <int M, bool FP32, bool MIX, bool GLOB>
__global__ __launch_bounds__(kFThreads) void k_force_cluster(
    int64_t nclusters, const float4 *__restrict__ pos32, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const int32_t *__restrict__ type, const double4 *__restrict__ c0,
    const int64_t *__restrict__ loff, const int2 *__restrict__ list, ForceArgs a, const double *__restrict__ tab64,
    double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz, double *__restrict__ pe,
    double *__restrict__ pv, int64_t *__restrict__ ph) {
  using real = typename std::conditional<FP32, float, double>::type;
  constexpr int NVAL = 3 * M + (GLOB ? 2 : 0);
  constexpr int NV = pow2_at_least(NVAL);
  __shared__ real stab[MIX ? 3 * kMaxTypes * kMaxTypes : 1];
  if (MIX) {
    for (int k = threadIdx.x; k < 3 * a.ntypes * a.ntypes; k += kFThreads) stab[k] = (real)tab64[k];
    __syncthreads();
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int ic = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave));
  double e_lane = 0., v_lane = 0.;
  int64_t hits = 0;
  if (ic < nclusters) {
    const double4 ci = c0[ic];
    const int nreal = (int)ci.w;
    real xi[M], yi[M], zi[M];
    int ti[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const int64_t s = (int64_t)ic * M + q;
      if constexpr (FP32) {
        const float4 p = pos32[s];
        xi[q] = (real)p.x; yi[q] = (real)p.y; zi[q] = (real)p.z;
        ti[q] = __float_as_int(p.w);
      } else {
        xi[q] = (real)x[s]; yi[q] = (real)y[s]; zi[q] = (real)z[s];
        ti[q] = type[s];
      }
    }
    real ax[M], ay[M], az[M], ae[M], av[M];
#pragma unroll
    for (int q = 0; q < M; ++q) { ax[q] = ay[q] = az[q] = ae[q] = av[q] = (real)0; }
    const int64_t beg = loff[ic];
    const int nj = (int)(loff[ic + 1] - beg) * M;
    const real rc2 = (real)a.rc2;
    for (int base = 0; base < nj; base += kWave) {
      const int k = base + lane;
      const bool valid = k < nj;
      const int kk = valid ? k : nj - 1;
      const int2 ent = list[beg + kk / M];
      const int sid = ent.y;
      const double shx = (double)(sid % 3 - 1) * a.L[0];
      const double shy = (double)((sid / 3) % 3 - 1) * a.L[1];
      const double shz = (double)(sid / 9 - 1) * a.L[2];
      const int64_t sj = (int64_t)ent.x * M + (kk & (M - 1));
      real xj, yj, zj;
      int tj;
      if constexpr (FP32) {
        const double4 cj = c0[ent.x];
        const float4 p = pos32[sj];
        xj = p.x + (float)((cj.x + shx) - ci.x);
        yj = p.y + (float)((cj.y + shy) - ci.y);
        zj = p.z + (float)((cj.z + shz) - ci.z);
        tj = __float_as_int(p.w);
      } else {
        xj = x[sj] + shx;
        yj = y[sj] + shy;
        zj = z[sj] + shz;
        tj = type[sj];
      }
#pragma unroll
      for (int q = 0; q < M; ++q) {
        if (q < nreal) {
          const real drx = xi[q] - xj, dry = yi[q] - yj, drz = zi[q] - zj;
          const real dr2 = drx * drx + dry * dry + drz * drz;
          real eps24, sig2, shift6;
          if constexpr (MIX) {
            const int t = 3 * (ti[q] * a.ntypes + tj);
            eps24 = stab[t]; sig2 = stab[t + 1]; shift6 = stab[t + 2];
          } else {
            eps24 = (real)a.eps24; sig2 = (real)a.sig2; shift6 = (real)a.shift6;
          }
          const bool mask = valid && dr2 <= rc2 && dr2 > (real)0;
          real fac, lj12m6;
          if constexpr (FP32) {
            const float inv = __builtin_amdgcn_rcpf(dr2);
            const float lj2 = sig2 * inv;
            const float lj6 = lj2 * lj2 * lj2;
            lj12m6 = lj6 * lj6 - lj6;
            fac = mask ? eps24 * (lj6 * lj6 + lj12m6) * inv : 0.f;
          } else {
            const double invdr2 = 1. / dr2;
            const double lj2 = sig2 * invdr2;
            const double lj6 = lj2 * lj2 * lj2;
            const double lj12 = lj6 * lj6;
            lj12m6 = lj12 - lj6;
            fac = mask ? eps24 * (lj12 + lj12m6) * invdr2 : 0.;
          }
          ax[q] += drx * fac;
          ay[q] += dry * fac;
          az[q] += drz * fac;
          if (GLOB) {
            ae[q] += mask ? eps24 * lj12m6 + shift6 : (real)0;
            av[q] += dr2 * fac;
          }
          hits += __popcll(__ballot(mask));
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
    if (GLOB) {
      double se = 0., sv = 0.;
#pragma unroll
      for (int q = 0; q < M; ++q) { se += (double)ae[q]; sv += (double)av[q]; }
      v[3 * M] = se;
      v[3 * M + 1] = sv;
    }
    wave_reduce_scatter<NV>(v, lane);
    const int q = vidx<NV>(lane);
    constexpr int group = 64 / NV;
    if ((lane & (group - 1)) == 0) {
      const int64_t s0 = (int64_t)ic * M;
      if (q < M) {
        if (q < nreal) fx[s0 + q] = v[0];
      } else if (q < 2 * M) {
        if (q - M < nreal) fy[s0 + q - M] = v[0];
      } else if (q < 3 * M) {
        if (q - 2 * M < nreal) fz[s0 + q - 2 * M] = v[0];
      } else if (GLOB && q == 3 * M) {
        e_lane = v[0];
      } else if (GLOB && q == 3 * M + 1) {
        v_lane = v[0];
      }
    }
  }
  block_partials(e_lane, v_lane, lane == 0 ? hits : 0, pe, pv, ph, GLOB);
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

void lc_force(Ctx &c, bool globals) {
  const ForceArgs a = make_args(c);
  const bool mix = c.lj.ntypes > 1;
  const double *tab = mix ? c.tab.p : nullptr;
  const int nb = blocks_for(c.n, kFThreads);
  c.part_e.ensure(nb);
  c.part_v.ensure(nb);
  c.part_hits.ensure(nb);
  State &s = c.st[c.cur];
  if (c.timing) LJG_HIP(hipEventRecord(c.ev0, c.stream));
  if (mix)
    hipLaunchKernelGGL(k_force_lc_fp64<true>, dim3(nb), dim3(kFThreads), 0, c.stream, c.n, s.x.p, s.y.p, s.z.p,
                       s.type.p, c.cell_of.p, c.cell_start.p, c.cg, a, tab, s.fx.p, s.fy.p, s.fz.p, c.part_e.p,
                       c.part_v.p, c.part_hits.p, (int32_t)globals);
  else
    hipLaunchKernelGGL(k_force_lc_fp64<false>, dim3(nb), dim3(kFThreads), 0, c.stream, c.n, s.x.p, s.y.p, s.z.p,
                       s.type.p, c.cell_of.p, c.cell_start.p, c.cg, a, tab, s.fx.p, s.fy.p, s.fz.p, c.part_e.p,
                       c.part_v.p, c.part_hits.p, (int32_t)globals);
  LJG_HIP(hipGetLastError());
  if (c.timing) LJG_HIP(hipEventRecord(c.ev1, c.stream));
}

template <int M, bool FP32>
static void launch_cluster(Ctx &c, bool mix, bool glob, int nb, const ForceArgs &a, const double *tab) {
  State &s = c.st[c.cur];
#define LJG_LAUNCH(MIXV, GLOBV)                                                                                    \
  hipLaunchKernelGGL((k_force_cluster<M, FP32, MIXV, GLOBV>), dim3(nb), dim3(kFThreads), 0, c.stream, c.nclusters, \
                     c.pos32.p, s.x.p, s.y.p, s.z.p, s.type.p, c.c0.p, c.list_off.p, c.list.p, a, tab, s.fx.p,     \
                     s.fy.p, s.fz.p, c.part_e.p, c.part_v.p, c.part_hits.p)
  if (mix) {
    if (glob) LJG_LAUNCH(true, true); else LJG_LAUNCH(true, false);
  } else {
    if (glob) LJG_LAUNCH(false, true); else LJG_LAUNCH(false, false);
  }
#undef LJG_LAUNCH
}

void cluster_force(Ctx &c, bool mixed, bool globals) {
  const ForceArgs a = make_args(c);
  const bool mix = c.lj.ntypes > 1;
  const double *tab = mix ? c.tab.p : nullptr;
  const int nb = blocks_for(c.nclusters, kWavesPerBlock);
  c.part_e.ensure(nb);
  c.part_v.ensure(nb);
  c.part_hits.ensure(nb);
  if (c.timing) LJG_HIP(hipEventRecord(c.ev0, c.stream));
  switch (c.M) {
    case 1: mixed ? launch_cluster<1, true>(c, mix, globals, nb, a, tab) : launch_cluster<1, false>(c, mix, globals, nb, a, tab); break;
    case 2: mixed ? launch_cluster<2, true>(c, mix, globals, nb, a, tab) : launch_cluster<2, false>(c, mix, globals, nb, a, tab); break;
    case 4: mixed ? launch_cluster<4, true>(c, mix, globals, nb, a, tab) : launch_cluster<4, false>(c, mix, globals, nb, a, tab); break;
    case 8: mixed ? launch_cluster<8, true>(c, mix, globals, nb, a, tab) : launch_cluster<8, false>(c, mix, globals, nb, a, tab); break;
    default: throw HipError{hipErrorInvalidValue, "cluster size must be 1, 2, 4 or 8", __LINE__};
  }
  LJG_HIP(hipGetLastError());
  if (c.timing) LJG_HIP(hipEventRecord(c.ev1, c.stream));
}

}  // namespace ljg
