// ljg_verlet.hip -- Verlet lists on the linked cells (AutoPas's VerletLists container and
// SoAFunctorVerlet, LJFunctor.h:569-578/812-1115; equivalently VerletClusterLists with cluster size 1),
// built and evaluated with 32-bit fixed-point coordinates.
//
// Positions are stored as X = round((x - box_min) * 2^32 / L) mod 2^32 per dimension.  The wrapping
// int32 difference Xi - Xj is then exactly the minimum-image separation in units of L / 2^32 (2.6e-8
// sigma for a 110-sigma box), with no rounding and no shift vectors, as long as every listed pair is
// closer than L/2 (boxes of at least 3 interaction lengths, which also keeps the 27 neighbour cells
// distinct).  Converting the difference to FP32 is the only rounding in a pair distance.
//
// Lists: one row of `cap` entries per particle, stored transposed per wavefront
// (entry k of lane l of wave w at ((w*cap + k)*64 + l)) so that the force kernel's list loads are
// coalesced; rows are padded up to the wave's longest list with the particle's own index (r^2 = 0,
// rejected by the cutoff test).  Compiled with -ffp-contract=off so the FP64 list criterion rounds as
// written; the force kernel's fused multiply-adds are explicit.
#include <cmath>

#include "ljg_internal.h"

namespace ljg {

constexpr int kVThreads = 256;

__device__ inline uint32_t to_fixed(double x, double mn, double scale_inv) {
  return (uint32_t)(uint64_t)__double2ll_rn((x - mn) * scale_inv);
}

// Padding slots (orig < 0) get type -1.
__global__ void k_to_fixed(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, const int32_t *__restrict__ type,
                           const int32_t *__restrict__ orig, FixGeom f, int4 *__restrict__ posfx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  posfx[i] = make_int4((int32_t)to_fixed(x[i], f.min[0], f.inv_scale[0]), (int32_t)to_fixed(y[i], f.min[1], f.inv_scale[1]),
                       (int32_t)to_fixed(z[i], f.min[2], f.inv_scale[2]), orig[i] < 0 ? -1 : type[i]);
}

void verlet_fixed_slots(Ctx &c, int64_t nslots) {
  State &s = c.st[c.cur];
  c.posfx.ensure(nslots);
  if (nslots <= 0) return;
  hipLaunchKernelGGL(k_to_fixed, dim3(blocks_for(nslots, 256)), dim3(256), 0, c.stream, nslots, s.x.p, s.y.p, s.z.p,
                     s.type.p, s.orig.p, c.fg, c.posfx.p);
  LJG_HIP(hipGetLastError());
}

void verlet_fixed(Ctx &c) { verlet_fixed_slots(c, c.n); }

void verlet_geometry(Ctx &c) {
  FixGeom &f = c.fg;
  bool cubic = true;
  for (int d = 0; d < 3; ++d) {
    const double L = c.prm.box_max[d] - c.prm.box_min[d];
    f.min[d] = c.prm.box_min[d];
    f.scale[d] = L / 4294967296.0;
    f.inv_scale[d] = 4294967296.0 / L;
    f.scale32[d] = (float)f.scale[d];
    cubic = cubic && L == c.prm.box_max[0] - c.prm.box_min[0];
  }
  f.cubic = cubic;
}

// Neighbour lists within ilen = rc + skin from the 27 cells around each particle's cell (cell order
// storage).  FP64 distance of the exact fixed-point separation, compared with ilen^2 (inclusive).
__global__ __launch_bounds__(kVThreads) void k_verlet_build(
    int64_t n, const int4 *__restrict__ posfx, const int32_t *__restrict__ cell_of,
    const int64_t *__restrict__ cell_start, CellGeom g, FixGeom f, double ilen2, int32_t cap,
    int32_t *__restrict__ count, int32_t *__restrict__ list) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w = i / kWave;
  const bool active = i < n;
  int32_t cnt = 0;
  if (active) {
    const int4 xi = posfx[i];
    const int32_t c = cell_of[i];
    const int cx = c % g.nh[0], cy = (c / g.nh[0]) % g.nh[1], cz = c / (g.nh[0] * g.nh[1]);
    const int no0 = g.nh[0] - 2, no1 = g.nh[1] - 2, no2 = g.nh[2] - 2;
    for (int oz = -1; oz <= 1; ++oz) {
      int jz = cz - 1 + oz;
      jz = jz < 0 ? jz + no2 : (jz >= no2 ? jz - no2 : jz);
      for (int oy = -1; oy <= 1; ++oy) {
        int jy = cy - 1 + oy;
        jy = jy < 0 ? jy + no1 : (jy >= no1 ? jy - no1 : jy);
        for (int ox = -1; ox <= 1; ++ox) {
          int jx = cx - 1 + ox;
          jx = jx < 0 ? jx + no0 : (jx >= no0 ? jx - no0 : jx);
          const int64_t cidx = ((int64_t)(jz + 1) * g.nh[1] + (jy + 1)) * g.nh[0] + (jx + 1);
          const int64_t sb = cell_start[cidx], se = cell_start[cidx + 1];
          for (int64_t j = sb; j < se; ++j) {
            if (j == i) continue;
            const int4 xj = posfx[j];
            const double dx = (double)(int32_t)((uint32_t)xi.x - (uint32_t)xj.x) * f.scale[0];
            const double dy = (double)(int32_t)((uint32_t)xi.y - (uint32_t)xj.y) * f.scale[1];
            const double dz = (double)(int32_t)((uint32_t)xi.z - (uint32_t)xj.z) * f.scale[2];
            if (dx * dx + dy * dy + dz * dz <= ilen2) {
              if (cnt < cap) list[(w * cap + cnt) * kWave + lane] = (int32_t)j;
              ++cnt;
            }
          }
        }
      }
    }
    count[i] = cnt;
  }
  // pad every row of the wave up to the wave's longest list (capped) with the particle itself
  int32_t wmax = cnt;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const int32_t o = __shfl_xor(wmax, off, kWave);
    wmax = o > wmax ? o : wmax;
  }
  const int32_t fill_to = wmax < cap ? wmax : cap;
  const int32_t self = active ? (int32_t)i : 0;
  for (int32_t k = cnt; k < fill_to; ++k) list[(w * cap + k) * kWave + lane] = self;
}

__global__ void k_zero_u64v(int64_t n, unsigned long long *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

void verlet_build(Ctx &c) {
  verlet_geometry(c);
  const FixGeom &f = c.fg;
  verlet_fixed(c);
  const int64_t n = c.n;
  const int64_t nwaves = (n + kWave - 1) / kWave;
  c.vcount.ensure(n);
  c.list_stats.ensure(2);
  const double ilen = c.prm.cutoff + c.prm.skin;
  int32_t cap = c.vcap > 0 ? c.vcap : 96;
  for (;;) {
    c.vlist.ensure((size_t)nwaves * cap * kWave);
    hipLaunchKernelGGL(k_zero_u64v, dim3(1), dim3(64), 0, c.stream, (int64_t)2, c.list_stats.p);
    LJG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_verlet_build, dim3(blocks_for(n, kVThreads)), dim3(kVThreads), 0, c.stream, n, c.posfx.p,
                       c.cell_of.p, c.cell_start.p, c.cg, f, ilen * ilen, cap, c.vcount.p, c.vlist.p);
    LJG_HIP(hipGetLastError());
    count_stats(c, c.vcount.p, n, c.list_stats.p);
    unsigned long long st[2] = {0, 0};
    LJG_HIP(hipMemcpyAsync(st, c.list_stats.p, sizeof st, hipMemcpyDeviceToHost, c.stream));
    LJG_HIP(hipStreamSynchronize(c.stream));
    const int32_t maxc = (int32_t)st[1];
    c.nentries = (int64_t)st[0];
    if (maxc <= cap) {
      const int32_t want = maxc + maxc / 8 + 8;
      c.vcap = want > cap ? want : (cap > 2 * want ? want : cap);
      break;
    }
    cap = maxc + maxc / 8 + 8;
  }
  c.vrow = cap;
  c.layout = Layout::Verlet;
}

// ---- force kernel ----------------------------------------------------------------------------------
// One lane per particle (cell order).  The cubic variant works in fixed-point units (r^2 and sigma^2 in
// units of (L/2^32)^2) and rescales the force by 2^32/L once per particle; the general variant scales
// each component.  Per-lane FP32 sums are flushed into FP64 accumulators every kFlush neighbours.
constexpr int kFlush = 16;

template <bool CUBIC, bool MIX, bool GLOB>
__global__ __launch_bounds__(kVThreads) void k_force_verlet(
    int64_t n, const int4 *__restrict__ posfx, int32_t cap, const int32_t *__restrict__ count,
    const int32_t *__restrict__ list, FixGeom f, uint32_t cut_lo, uint32_t cut_hi, double rc2d, float eps24_1,
    float sig2_1, float shift6_1,
    int32_t ntypes, const double *__restrict__ tab64, double *__restrict__ fx, double *__restrict__ fy,
    double *__restrict__ fz, double *__restrict__ pe, double *__restrict__ pv, int64_t *__restrict__ ph,
    unsigned long long *__restrict__ acc, SkipIf sk, Owned own) {
  if (skip_launch(sk)) return;
  __shared__ float stab[MIX ? 3 * kMaxTypes * kMaxTypes : 1];
  if constexpr (MIX) {
    // sigma^2 in the units the distances are computed in
    const float s2u = CUBIC ? (float)(f.inv_scale[0] * f.inv_scale[0]) : 1.f;
    for (int k = threadIdx.x; k < 3 * ntypes * ntypes; k += kVThreads)
      stab[k] = (float)tab64[k] * (k % 3 == 1 ? s2u : 1.f);
    __syncthreads();
  }
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w = i / kWave;
  const bool active = i < n;
  int4 xi = posfx[active ? i : 0];
  if (!active) {  // half a box away from particle 0 (which pads its rows): every test fails
    xi.x = (int32_t)((uint32_t)xi.x + 0x80000000u);
    xi.y = (int32_t)((uint32_t)xi.y + 0x80000000u);
    xi.z = (int32_t)((uint32_t)xi.z + 0x80000000u);
  }
  const int32_t myc = active ? count[i] : 0;
  int32_t nmax = myc;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const int32_t o = __shfl_xor(nmax, off, kWave);
    nmax = o > nmax ? o : nmax;
  }
  nmax = nmax < cap ? nmax : cap;
  const int ti = xi.w;
  double dax = 0., day = 0., daz = 0., dae = 0., dav = 0.;
  int hits = 0;
  const int32_t *row = list + w * (int64_t)cap * kWave + lane;
  int32_t jn = nmax > 0 ? row[0] : 0;
  int4 xjn = posfx[jn];
  for (int kb = 0; kb < nmax; kb += kFlush) {
    float ax = 0.f, ay = 0.f, az = 0.f, ae = 0.f, av = 0.f;
    const int ke = kb + kFlush < nmax ? kb + kFlush : nmax;
    for (int k = kb; k < ke; ++k) {
      const int4 xj = xjn;
      if (k + 1 < nmax) {
        jn = row[(int64_t)(k + 1) * kWave];
        xjn = posfx[jn];
      }
      float dx = (float)(int32_t)((uint32_t)xi.x - (uint32_t)xj.x);
      float dy = (float)(int32_t)((uint32_t)xi.y - (uint32_t)xj.y);
      float dz = (float)(int32_t)((uint32_t)xi.z - (uint32_t)xj.z);
      if constexpr (!CUBIC) {
        dx *= f.scale32[0];
        dy *= f.scale32[1];
        dz *= f.scale32[2];
      }
      const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
      // the cutoff decision is the FP64 one: pairs whose FP32 r^2 lies between cut_lo and cut_hi (rare)
      // are decided again by the FP64 r^2 of the exact fixed-point separation
      const uint32_t ub = __float_as_uint(r2) - 1u;  // r2 = +0 (the row's padding is the particle itself) wraps
      bool m = ub < cut_lo;
      if ((__builtin_amdgcn_ballot_w64(ub < cut_hi) & ~__builtin_amdgcn_ballot_w64(m)) != 0ull) {
        const double ex = (double)(int32_t)((uint32_t)xi.x - (uint32_t)xj.x) * f.scale[0];
        const double ey = (double)(int32_t)((uint32_t)xi.y - (uint32_t)xj.y) * f.scale[1];
        const double ez = (double)(int32_t)((uint32_t)xi.z - (uint32_t)xj.z) * f.scale[2];
        const double d2 = ex * ex + ey * ey + ez * ez;
        if (!m && ub < cut_hi) m = d2 > 0.0 && d2 <= rc2d;
      }
      float eps24 = eps24_1, sig2 = sig2_1, shift6 = shift6_1;
      if constexpr (MIX) {
        const int t = 3 * (ti * ntypes + xj.w);
        eps24 = stab[t]; sig2 = stab[t + 1]; shift6 = stab[t + 2];
      }
      const float inv = __builtin_amdgcn_rcpf(r2);
      const float lj2 = sig2 * inv;
      const float lj6 = lj2 * lj2 * lj2;
      const float fv = (lj6 * inv) * __builtin_fmaf(eps24 + eps24, lj6, -eps24);
      const float fac = m ? fv : 0.f;
      ax = __builtin_fmaf(dx, fac, ax);
      ay = __builtin_fmaf(dy, fac, ay);
      az = __builtin_fmaf(dz, fac, az);
      hits += m ? 1 : 0;
      if constexpr (GLOB) {
        ae += m ? __builtin_fmaf(eps24, lj6 * lj6 - lj6, shift6) : 0.f;
        av = __builtin_fmaf(m ? r2 : 0.f, fac, av);
      }
    }
    dax += (double)ax;
    day += (double)ay;
    daz += (double)az;
    if constexpr (GLOB) {
      dae += (double)ae;
      dav += (double)av;
    }
  }
  if (active) {
    // cubic: forces were accumulated in fixed-point units; F = sum(dx_u * fac_u) * (2^32 / L)
    const double k = CUBIC ? f.inv_scale[0] : 1.0;
    fx[i] = dax * k;
    fy[i] = day * k;
    fz[i] = daz * k;
  }
  // per-block partials (fixed tree) and the pair count
  __shared__ double se[kVThreads], sv[kVThreads];
  __shared__ int64_t sh[kVThreads];
  const bool counted = GLOB && active && is_owned(own, i);
  se[threadIdx.x] = counted ? dae : 0.;
  sv[threadIdx.x] = counted ? dav : 0.;
  sh[threadIdx.x] = active ? hits : 0;
  __syncthreads();
  for (int s = kVThreads / 2; s > 0; s >>= 1) {
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

void verlet_force(Ctx &c, bool globals) {
  const bool mix = c.lj.ntypes > 1;
  const double *tab = mix ? c.tab.p : nullptr;
  const int nb = blocks_for(c.n, kVThreads);
  c.part_e.ensure(nb);
  c.part_v.ensure(nb);
  c.part_hits.ensure(nb);
  c.part_items = nb;
  const FixGeom &f = c.fg;
  const double u2 = f.cubic ? f.inv_scale[0] * f.inv_scale[0] : 1.0;  // (2^32/L)^2 in the cubic variant
  const float rc2 = (float)(c.lj.rc2 * u2);
  const float sig2 = (float)(c.lj.sig2[0] * u2);
  // r^2 from the exact integer separation: one conversion, one scale and the sums, ~9 roundings
  uint32_t cut_lo, cut_hi;
  cutoff_band(rc2, std::ldexp(1.0, -18), cut_lo, cut_hi);
  State &s = c.st[c.cur];
  if (c.timing) LJG_HIP(hipEventRecord(c.ev0, c.stream));
#define LJG_LAUNCH(CUB, MIXV, GLOBV)                                                                             \
  hipLaunchKernelGGL((k_force_verlet<CUB, MIXV, GLOBV>), dim3(nb), dim3(kVThreads), 0, c.stream, c.n, c.posfx.p, \
                     c.vrow, c.vcount.p, c.vlist.p, f, cut_lo, cut_hi, c.lj.rc2, (float)c.lj.eps24[0], sig2,      \
                     (float)c.lj.shift6[0],                                                                       \
                     c.lj.ntypes, tab, s.fx.p, s.fy.p, s.fz.p, c.part_e.p, c.part_v.p, c.part_hits.p, c.hits_acc, c.skip, \
                     owned_of(c))
#define LJG_LAUNCH2(CUB)                                                 \
  if (mix) {                                                             \
    if (globals) LJG_LAUNCH(CUB, true, true); else LJG_LAUNCH(CUB, true, false); \
  } else {                                                               \
    if (globals) LJG_LAUNCH(CUB, false, true); else LJG_LAUNCH(CUB, false, false); \
  }
  if (f.cubic) {
    LJG_LAUNCH2(true)
  } else {
    LJG_LAUNCH2(false)
  }
#undef LJG_LAUNCH2
#undef LJG_LAUNCH
  LJG_HIP(hipGetLastError());
  if (c.timing) LJG_HIP(hipEventRecord(c.ev1, c.stream));
}

}  // namespace ljg
