// ljg_md.hip -- device-wide scan, state permutation, velocity-Verlet integrator, periodic wrap and the
// deterministic reductions of per-work-item partial sums.
#include <cmath>

#include <chrono>

#include "ljg_internal.h"

namespace ljg {

// ---------------------------------------------------------------------------------------------------
// Exclusive scan: out[i] = sum_{k<i} in[k] for i in [0, n]; out[n] is the total.  Three phases (block
// scan, scan of block totals, add), recursive on the block totals.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <class TIn>
__global__ __launch_bounds__(kScanThreads) void k_scan_local(const TIn *__restrict__ in, int64_t n_in,
                                                              int64_t n_out, int64_t *__restrict__ out,
                                                              int64_t *__restrict__ block_tot) {
  __shared__ int64_t warp_tot[kScanThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    v[k] = i < n_in ? (int64_t)in[i] : 0;
    s += v[k];
  }
  // inclusive scan of per-thread sums within the wave
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  int64_t incl = s;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int64_t t = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += t;
  }
  if (lane == kWave - 1) warp_tot[w] = incl;
  __syncthreads();
  int64_t wprefix = 0;
  for (int k = 0; k < w; ++k) wprefix += warp_tot[k];
  int64_t run = wprefix + incl - s;  // exclusive prefix of this thread inside the block
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n_out) out[i] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanThreads - 1) block_tot[blockIdx.x] = wprefix + incl;
}

__global__ void k_scan_add(int64_t *__restrict__ out, int64_t n_out, const int64_t *__restrict__ block_off) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_out) out[i] += block_off[i / kScanTile];
}

template <class TIn>
static void scan_impl(hipStream_t st, const TIn *in, int64_t n_in, int64_t *out, int64_t n_out, int64_t *scratch,
                      int64_t scratch_len) {
  const int64_t nb = (n_out + kScanTile - 1) / kScanTile;
  if (scratch_len < 2 * (nb + 1)) throw HipError{hipErrorInvalidValue, "scan scratch too small", __LINE__};
  int64_t *tot = scratch, *tot_scan = scratch + (nb + 1);
  hipLaunchKernelGGL(k_scan_local<TIn>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n_in, n_out, out, tot);
  LJG_HIP(hipGetLastError());
  if (nb > 1) {
    scan_impl<int64_t>(st, tot, nb, tot_scan, nb, scratch + 2 * (nb + 1), scratch_len - 2 * (nb + 1));
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n_out, 256)), dim3(256), 0, st, out, n_out, tot_scan);
    LJG_HIP(hipGetLastError());
  }
}

int64_t scan_tmp_len(int64_t n) {
  int64_t need = 0, m = n + 1;
  do {
    const int64_t nb = (m + kScanTile - 1) / kScanTile;
    need += 2 * (nb + 1);
    m = nb;
  } while (m > 1);
  return need + 8;
}

void exclusive_scan_i32_to_i64_on(hipStream_t st, const int32_t *in, int64_t *out, int64_t n, int64_t *tmp,
                                  int64_t tmp_len) {
  scan_impl<int32_t>(st, in, n, out, n + 1, tmp, tmp_len);
}

void exclusive_scan_i32_to_i64(Ctx &c, const int32_t *in, int64_t *out, int64_t n) {
  c.scan_tmp.ensure((size_t)scan_tmp_len(n));
  scan_impl<int32_t>(c.stream, in, n, out, n + 1, c.scan_tmp.p, (int64_t)c.scan_tmp.cap);
}

// ---------------------------------------------------------------------------------------------------
// State permutation: new slot s takes the particle of old slot src[s]; src[s] < 0 makes a padding slot.
// Every permutation is a rebuild, so the rebuild reference positions are set to the positions here, and
// when posfx is given the slots' 32-bit fixed-point positions are written too (type -1 for padding).
// Inside the step loop the force is recomputed right after a rebuild, so it is not carried.
__global__ void k_permute(int64_t nnew, const int32_t *__restrict__ src, StateView o, StateView n,
                          int32_t carry_force, int4 *__restrict__ posfx, FixGeom f) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnew) return;
  const int32_t k = src[s];
  if (k >= 0) {
    const double x = o.x[k], y = o.y[k], z = o.z[k];
    const int32_t t = o.type[k];
    n.x[s] = x;   n.y[s] = y;   n.z[s] = z;
    n.rx[s] = x;  n.ry[s] = y;  n.rz[s] = z;
    n.vx[s] = o.vx[k]; n.vy[s] = o.vy[k]; n.vz[s] = o.vz[k];
    if (carry_force) { n.fx[s] = o.fx[k]; n.fy[s] = o.fy[k]; n.fz[s] = o.fz[k]; }
    n.ofx[s] = o.ofx[k]; n.ofy[s] = o.ofy[k]; n.ofz[s] = o.ofz[k];
    n.type[s] = t; n.orig[s] = o.orig[k]; n.id[s] = o.id[k];
    if (posfx)
      posfx[s] = make_int4((int32_t)(uint32_t)(uint64_t)__double2ll_rn((x - f.min[0]) * f.inv_scale[0]),
                           (int32_t)(uint32_t)(uint64_t)__double2ll_rn((y - f.min[1]) * f.inv_scale[1]),
                           (int32_t)(uint32_t)(uint64_t)__double2ll_rn((z - f.min[2]) * f.inv_scale[2]), t);
  } else {
    n.x[s] = kDummyPos; n.y[s] = kDummyPos; n.z[s] = kDummyPos;
    n.vx[s] = 0.; n.vy[s] = 0.; n.vz[s] = 0.;
    n.fx[s] = 0.; n.fy[s] = 0.; n.fz[s] = 0.;
    n.ofx[s] = 0.; n.ofy[s] = 0.; n.ofz[s] = 0.;
    n.rx[s] = kDummyPos; n.ry[s] = kDummyPos; n.rz[s] = kDummyPos;
    n.type[s] = 0; n.orig[s] = -1; n.id[s] = -1;
    if (posfx) posfx[s] = make_int4(0, 0, 0, -1);
  }
}

// The same gather over the slots of the marked ensemble boxes only, into the other buffer at the same
// slots, then copied back: each box's particles stay in the box's region, so the gather reads no slot
// another thread writes, and the unmarked boxes' slots are never touched.
__global__ void k_permute_boxes(int64_t nitems, Segs sg, int32_t M, const int32_t *__restrict__ src,
                                const int32_t *__restrict__ cluster_box, const int32_t *__restrict__ box_active,
                                StateView o, StateView n, int32_t carry_force,
                                const unsigned long long *__restrict__ skip) {
  const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nitems || (skip && *skip)) return;
  const int64_t s = seg_index(sg, it);
  if (box_active[cluster_box[s / M]] == 0) return;
  const int32_t k = src[s];
  if (k >= 0) {
    const double x = o.x[k], y = o.y[k], z = o.z[k];
    n.x[s] = x;   n.y[s] = y;   n.z[s] = z;
    n.rx[s] = x;  n.ry[s] = y;  n.rz[s] = z;
    n.vx[s] = o.vx[k]; n.vy[s] = o.vy[k]; n.vz[s] = o.vz[k];
    if (carry_force) { n.fx[s] = o.fx[k]; n.fy[s] = o.fy[k]; n.fz[s] = o.fz[k]; }
    else { n.fx[s] = 0.; n.fy[s] = 0.; n.fz[s] = 0.; }
    n.ofx[s] = o.ofx[k]; n.ofy[s] = o.ofy[k]; n.ofz[s] = o.ofz[k];
    n.type[s] = o.type[k]; n.orig[s] = o.orig[k]; n.id[s] = o.id[k];
  } else {
    n.x[s] = kDummyPos; n.y[s] = kDummyPos; n.z[s] = kDummyPos;
    n.vx[s] = 0.; n.vy[s] = 0.; n.vz[s] = 0.;
    n.fx[s] = 0.; n.fy[s] = 0.; n.fz[s] = 0.;
    n.ofx[s] = 0.; n.ofy[s] = 0.; n.ofz[s] = 0.;
    n.rx[s] = kDummyPos; n.ry[s] = kDummyPos; n.rz[s] = kDummyPos;
    n.type[s] = 0; n.orig[s] = -1; n.id[s] = -1;
  }
}

__global__ void k_copy_boxes(int64_t nitems, Segs sg, int32_t M, const int32_t *__restrict__ cluster_box,
                             const int32_t *__restrict__ box_active, StateView from, StateView to,
                             const unsigned long long *__restrict__ skip) {
  const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nitems || (skip && *skip)) return;
  const int64_t s = seg_index(sg, it);
  if (box_active[cluster_box[s / M]] == 0) return;
  to.x[s] = from.x[s];   to.y[s] = from.y[s];   to.z[s] = from.z[s];
  to.vx[s] = from.vx[s]; to.vy[s] = from.vy[s]; to.vz[s] = from.vz[s];
  to.fx[s] = from.fx[s]; to.fy[s] = from.fy[s]; to.fz[s] = from.fz[s];
  to.ofx[s] = from.ofx[s]; to.ofy[s] = from.ofy[s]; to.ofz[s] = from.ofz[s];
  to.rx[s] = from.rx[s]; to.ry[s] = from.ry[s]; to.rz[s] = from.rz[s];
  to.type[s] = from.type[s]; to.orig[s] = from.orig[s]; to.id[s] = from.id[s];
}

void permute_active(Ctx &c, const int32_t *src, const int32_t *active, const unsigned long long *skip) {
  const int other = 1 - c.cur;
  c.st[other].ensure((size_t)c.nslots);
  const int64_t ni = c.seg_nslots;
  if (ni == 0) return;
  hipLaunchKernelGGL(k_permute_boxes, dim3(blocks_for(ni, 256)), dim3(256), 0, c.stream, ni, c.seg_slots, c.M, src,
                     c.cluster_box.p, active, view(c.st[c.cur]), view(c.st[other]), (int32_t)c.carry_force, skip);
  LJG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_copy_boxes, dim3(blocks_for(ni, 256)), dim3(256), 0, c.stream, ni, c.seg_slots, c.M,
                     c.cluster_box.p, active, view(c.st[other]), view(c.st[c.cur]), skip);
  LJG_HIP(hipGetLastError());
}

void permute_state(Ctx &c, const int32_t *src, int64_t new_slots, int4 *posfx) {
  const int other = 1 - c.cur;
  c.st[other].ensure((size_t)new_slots);
  hipLaunchKernelGGL(k_permute, dim3(blocks_for(new_slots, 256)), dim3(256), 0, c.stream, new_slots, src,
                     view(c.st[c.cur]), view(c.st[other]), (int32_t)c.carry_force, posfx, c.fg);
  LJG_HIP(hipGetLastError());
  c.cur = other;
  c.nslots = new_slots;
}

// ---------------------------------------------------------------------------------------------------
// Velocity Verlet, TimeDiscretization.cpp:152-165 (kick: v += (F + oldF) * (dt/(2m))) followed by
// TimeDiscretization.cpp:17-69 (drift: oldF = F; x += v*dt + F*(dt*dt/(2m))).  Fusing the kick of
// step k with the drift of step k+1 keeps each particle's FP64 operation sequence unchanged.  The
// drift also refreshes the FP32 cluster-relative positions and raises the displacement flag of
// LogicHandler.h:1000-1015 when |x - x_rebuild|^2 >= (skin/2)^2.
struct MdArgs {
  double dt;
  double half_skin2;
  int32_t kick, drift, write_pos32, write_fixed, M, stamp;
  int32_t *box_flag;            // ensemble: set to 1 for the box whose particle raised the flag (else null)
  const int32_t *cluster_box;   // ensemble: box of each cluster
  FixGeom fg;
  double kick_fac[kMaxTypes];   // dt / (2m), as TimeDiscretization.cpp:160 forms it
  double drift_fac[kMaxTypes];  // dt * dt / (2m), as TimeDiscretization.cpp:40 forms it
};

__global__ void k_kick_drift(int64_t nslots, StateView s, MdArgs a, const double4 *__restrict__ c0,
                             float4 *__restrict__ pos32, double4 *__restrict__ posd, int4 *__restrict__ posfx,
                             int32_t *__restrict__ flag, int32_t *__restrict__ hflag, SkipIf sk) {
  if (skip_launch(sk)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  if (s.orig[i] < 0) return;
  const int t = s.type[i];
  double fx = s.fx[i], fy = s.fy[i], fz = s.fz[i];
  double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
  if (a.kick) {
    const double fac = a.kick_fac[t];
    vx += (fx + s.ofx[i]) * fac;
    vy += (fy + s.ofy[i]) * fac;
    vz += (fz + s.ofz[i]) * fac;
    s.vx[i] = vx; s.vy[i] = vy; s.vz[i] = vz;
  }
  if (!a.drift) return;
  const double ff = a.drift_fac[t];
  const double dx = vx * a.dt + fx * ff;
  const double dy = vy * a.dt + fy * ff;
  const double dz = vz * a.dt + fz * ff;
  s.ofx[i] = fx; s.ofy[i] = fy; s.ofz[i] = fz;
  const double x = s.x[i] + dx, y = s.y[i] + dy, z = s.z[i] + dz;
  s.x[i] = x; s.y[i] = y; s.z[i] = z;
  const double ex = x - s.rx[i], ey = y - s.ry[i], ez = z - s.rz[i];
  if (ex * ex + ey * ey + ez * ez >= a.half_skin2) {
    flag[0] = a.stamp;
    hflag[0] = a.stamp;  // pinned host copy, read by the step loop without a device-to-host copy
    if (a.box_flag) a.box_flag[a.cluster_box[i / a.M]] = 1;
  }
  if (a.write_pos32) {
    const double4 r = c0[i / a.M];  // padding slots returned above and keep their sentinel positions
    pos32[i] = make_float4((float)(x - r.x), (float)(y - r.y), (float)(z - r.z), __int_as_float(t));
    posd[i] = make_double4(x, y, z, 0.);
  }
  if (a.write_fixed) {
    posfx[i] = make_int4((int32_t)(uint32_t)(uint64_t)__double2ll_rn((x - a.fg.min[0]) * a.fg.inv_scale[0]),
                         (int32_t)(uint32_t)(uint64_t)__double2ll_rn((y - a.fg.min[1]) * a.fg.inv_scale[1]),
                         (int32_t)(uint32_t)(uint64_t)__double2ll_rn((z - a.fg.min[2]) * a.fg.inv_scale[2]), t);
  }
}

void md_kick_drift(Ctx &c, bool kick, bool drift, double dt, int32_t stamp, SkipIf skip) {
  MdArgs a{};
  a.dt = dt;
  a.stamp = stamp;
  const double hs = c.prm.skin * 0.5;
  a.half_skin2 = hs * hs;
  a.kick = kick;
  a.drift = drift;
  a.write_pos32 = c.layout == Layout::Clusters && !c.jrows_valid;
  a.write_fixed = c.layout == Layout::Verlet || (c.layout == Layout::Clusters && c.jrows_valid);
  a.fg = c.fg;
  a.M = c.M;
  if (c.nboxes > 0 && c.regions_valid && c.layout == Layout::Clusters) {
    a.box_flag = c.box_flag.p;
    a.cluster_box = c.cluster_box.p;
  }
  for (int t = 0; t < kMaxTypes; ++t) {
    const double m = c.lj.mass[t] > 0. ? c.lj.mass[t] : 1.;
    a.kick_fac[t] = dt / (2 * m);
    a.drift_fac[t] = dt * dt / (2 * m);
  }
  hipLaunchKernelGGL(k_kick_drift, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots,
                     view(c.st[c.cur]), a, c.c0.p, c.pos32.p, c.posd.p, c.posfx.p, c.flag.p, c.host_flag_dev, skip);
  LJG_HIP(hipGetLastError());
}

// Positions written by the caller between two force evaluations (upload order in, storage order out),
// with the writes of k_kick_drift's drift: the displacement check against the rebuild positions and the
// FP32 cluster-relative / fixed-point copies of the layout.
__global__ void k_set_positions(int64_t nslots, StateView s, int64_t n, const double *__restrict__ pos,
                                const double *__restrict__ vel, MdArgs a, const double4 *__restrict__ c0,
                                float4 *__restrict__ pos32, double4 *__restrict__ posd, unsigned long long *ext,
                                int4 *__restrict__ posfx, int32_t *__restrict__ moved) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t o = i < nslots ? s.orig[i] : -1;
  float e = 0.f;
  if (o >= 0) {
    const double x = pos[o], y = pos[n + o], z = pos[2 * n + o];
    s.x[i] = x; s.y[i] = y; s.z[i] = z;
    if (vel) {
      s.vx[i] = vel[o]; s.vy[i] = vel[n + o]; s.vz[i] = vel[2 * n + o];
    }
    const double ex = x - s.rx[i], ey = y - s.ry[i], ez = z - s.rz[i];
    if (ex * ex + ey * ey + ez * ez >= a.half_skin2) moved[0] = 1;
    const int t = s.type[i];
    if (a.write_pos32) {
      const double4 r = c0[i / a.M];
      const float4 p = make_float4((float)(x - r.x), (float)(y - r.y), (float)(z - r.z), __int_as_float(t));
      pos32[i] = p;
      posd[i] = make_double4(x, y, z, 0.);
      e = fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z)));
    }
    if (a.write_fixed) {
      posfx[i] = make_int4((int32_t)(uint32_t)(uint64_t)__double2ll_rn((x - a.fg.min[0]) * a.fg.inv_scale[0]),
                           (int32_t)(uint32_t)(uint64_t)__double2ll_rn((y - a.fg.min[1]) * a.fg.inv_scale[1]),
                           (int32_t)(uint32_t)(uint64_t)__double2ll_rn((z - a.fg.min[2]) * a.fg.inv_scale[2]), t);
    }
  }
  // the caller may move particles farther than the step loop does: the cutoff band of the FP32 cluster
  // kernel follows the largest cluster-relative offset
  if (a.write_pos32) block_max_to_lines(e, ext);
}

void md_set_positions(Ctx &c, const double *pos, const double *vel) {
  MdArgs a{};
  const double hs = c.prm.skin * 0.5;
  a.half_skin2 = hs * hs;
  a.write_pos32 = c.layout == Layout::Clusters && !c.jrows_valid;
  a.write_fixed = c.layout == Layout::Verlet || (c.layout == Layout::Clusters && c.jrows_valid);
  a.fg = c.fg;
  a.M = c.M;
  hipLaunchKernelGGL(k_set_positions, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots,
                     view(c.st[c.cur]), c.n, pos, vel, a, c.c0.p, c.pos32.p, c.posd.p, c.pos32_ext.p, c.posfx.p,
                     c.flag.p + 2);
  LJG_HIP(hipGetLastError());
}

IntegArgs integ_args(Ctx &c, bool kick, bool drift, double dt, int32_t stamp) {
  State &s = c.st[c.cur];
  IntegArgs a{};
  a.vx = s.vx.p; a.vy = s.vy.p; a.vz = s.vz.p;
  a.ofx = s.ofx.p; a.ofy = s.ofy.p; a.ofz = s.ofz.p;
  a.x = s.x.p; a.y = s.y.p; a.z = s.z.p;
  a.rx = s.rx.p; a.ry = s.ry.p; a.rz = s.rz.p;
  if (c.jrows_valid) {
    c.posfx_next.ensure(c.nslots);
    a.posfx_next = c.posfx_next.p;
  } else {
    c.pos32_next.ensure(c.nslots);
    a.pos32_next = c.pos32_next.p;
    c.posd_next.ensure(c.nslots);
    a.posd_next = c.posd_next.p;
  }
  a.c0 = c.c0.p;
  a.flag = c.flag.p;
  a.hflag = c.host_flag_dev;
  if (c.nboxes > 0 && c.regions_valid) {
    a.box_flag = c.box_flag.p;
    a.cluster_box = c.cluster_box.p;
  }
  a.dt = dt;
  const double hs = c.prm.skin * 0.5;
  a.half_skin2 = hs * hs;
  a.kick = kick;
  a.drift = drift;
  a.stamp = stamp;
  for (int t = 0; t < kMaxTypes; ++t) {
    const double m = c.lj.mass[t] > 0. ? c.lj.mass[t] : 1.;
    a.kick_fac[t] = dt / (2 * m);
    a.drift_fac[t] = dt * dt / (2 * m);
  }
  return a;
}


// Periodic wrap of particles that left the box, RegularGridDecomposition.cpp:563-587.
__global__ void k_wrap(int64_t nslots, StateView s, WrapArgs w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots || s.orig[i] < 0) return;
  double *c[3] = {s.x, s.y, s.z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double p = c[d][i], q = wrap_coord(p, d, w);
    if (q != p) c[d][i] = q;
  }
}

WrapArgs wrap_args(const double lo[3], const double hi[3]) {
  WrapArgs w{};
  for (int d = 0; d < 3; ++d) {
    w.mn[d] = lo[d];
    w.mx[d] = hi[d];
    w.L[d] = hi[d] - lo[d];
    w.just_inside[d] = std::nextafter(hi[d], lo[d]);
  }
  return w;
}

WrapArgs wrap_args(const ljg_params &p) { return wrap_args(p.box_min, p.box_max); }

void wrap_positions(Ctx &c) {
  hipLaunchKernelGGL(k_wrap, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots,
                     view(c.st[c.cur]), wrap_args(c.prm));
  LJG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------------
// Deterministic reduction of per-work-item partials: one workgroup, fixed assignment of items to
// threads and a fixed tree, so the result does not depend on scheduling.
__global__ __launch_bounds__(1024) void k_reduce(int64_t n, const double *__restrict__ e,
                                                 const double *__restrict__ v, const int64_t *__restrict__ h,
                                                 double *__restrict__ out, int64_t *__restrict__ out_h,
                                                 int32_t globals) {
  __shared__ double se[1024], sv[1024];
  __shared__ int64_t sh[1024];
  double ae = 0., av = 0.;
  int64_t ah = 0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) {
    if (globals) {
      ae += e[i];
      av += v[i];
    }
    ah += h[i];
  }
  se[threadIdx.x] = ae;
  sv[threadIdx.x] = av;
  sh[threadIdx.x] = ah;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sv[threadIdx.x] += sv[threadIdx.x + s];
      sh[threadIdx.x] += sh[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = se[0];
    out[1] = sv[0];
    out_h[0] = sh[0];
  }
}

constexpr int kStatThreads = 256;

__global__ __launch_bounds__(kStatThreads) void k_count_stats(int64_t n, const int32_t *__restrict__ cnt,
                                                              unsigned long long *__restrict__ stats) {
  __shared__ unsigned long long ss[kStatThreads];
  __shared__ int32_t sm[kStatThreads];
  unsigned long long sum = 0;
  int32_t mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * kStatThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kStatThreads) {
    const int32_t v = cnt[i];
    sum += (unsigned long long)v;
    mx = v > mx ? v : mx;
  }
  ss[threadIdx.x] = sum;
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int k = kStatThreads / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      ss[threadIdx.x] += ss[threadIdx.x + k];
      sm[threadIdx.x] = sm[threadIdx.x + k] > sm[threadIdx.x] ? sm[threadIdx.x + k] : sm[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&stats[0], ss[0]);
    atomicMax(&stats[1], (unsigned long long)sm[0]);
  }
}

// Copies n words of src, then n2 words of src2, behind the marker word, and writes the marker last.
__global__ void k_mark(unsigned long long *__restrict__ dst, const unsigned long long *__restrict__ src, int32_t n,
                       const unsigned long long *__restrict__ src2, int32_t n2, unsigned long long seq) {
  const int32_t i = (int32_t)threadIdx.x;
  if (i < n) dst[1 + i] = src[i];
  else if (i < n + n2) dst[1 + i] = src2[i - n];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();
    dst[0] = seq;
  }
}

static unsigned long long mark_launch(Ctx &c, const unsigned long long *src, int n, const unsigned long long *src2,
                                      int n2) {
  if (n < 0 || n2 < 0 || n + n2 > kMarkWords)
    throw HipError{hipErrorInvalidValue, "stream_wait reads at most kMarkWords words", __LINE__};
  static_assert(kMarkWords < 64, "k_mark copies a word per lane of one wavefront");
  const unsigned long long seq = ++c.mark_seq;
  hipLaunchKernelGGL(k_mark, dim3(1), dim3(64), 0, c.stream, c.pin_mark_dev, src, (int32_t)n, src2, (int32_t)n2, seq);
  LJG_HIP(hipGetLastError());
  return seq;
}

void stream_spin(Ctx &c, unsigned long long seq, int n, unsigned long long *out) {
  if (seq != c.mark_seq) throw HipError{hipErrorInvalidValue, "another stream marker was queued", __LINE__};
  volatile unsigned long long *m = c.pin_mark;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint64_t spins = 0; m[0] != seq; ++spins) {
    if ((spins & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
      LJG_HIP(hipStreamSynchronize(c.stream));
      if (m[0] != seq) throw HipError{hipErrorUnknown, "stream marker not written", __LINE__};
      break;
    }
  }
  for (int i = 0; i < n; ++i) out[i] = m[1 + i];
}

unsigned long long stream_mark(Ctx &c, const unsigned long long *src, int n) {
  return mark_launch(c, src, n, nullptr, 0);
}

void stream_wait(Ctx &c, const unsigned long long *src, int n, unsigned long long *out,
                 const unsigned long long *src2, int n2) {
  const unsigned long long seq = mark_launch(c, src, n, src2, n2);
  stream_spin(c, seq, n + n2, out);
}

void count_stats(Ctx &c, const int32_t *cnt, int64_t n, unsigned long long *stats) {
  if (n <= 0) return;
  int nb = blocks_for(n, kStatThreads * 16);
  nb = nb < 256 ? nb : 256;
  hipLaunchKernelGGL(k_count_stats, dim3(nb), dim3(kStatThreads), 0, c.stream, n, cnt, stats);
  LJG_HIP(hipGetLastError());
}

void reduce_launch(hipStream_t st, int64_t nitems, const double *pe, const double *pv, const int64_t *ph,
                   double *out, int64_t *out_h, bool globals) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, nitems, pe, pv, ph, out, out_h, (int32_t)globals);
  LJG_HIP(hipGetLastError());
}

void reduce_partials(Ctx &c, int64_t nitems, ljg_globals *out, bool globals) {
  c.red_out.ensure(2);
  c.red_hits.ensure(1);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c.stream, nitems, c.part_e.p, c.part_v.p, c.part_hits.p,
                     c.red_out.p, c.red_hits.p, (int32_t)globals);
  LJG_HIP(hipGetLastError());
  double hv[2] = {0., 0.};
  int64_t hh = 0;
  LJG_HIP(hipMemcpyAsync(hv, c.red_out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  LJG_HIP(hipMemcpyAsync(&hh, c.red_hits.p, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  LJG_HIP(hipStreamSynchronize(c.stream));
  // Full lists visit each unique pair from both sides: LJFunctor::endTraversal halves the sums and
  // divides the energy by 6 (LJFunctor.h:661-685).
  out->potential_energy = hv[0] * 0.5 / 6.;
  out->virial = hv[1] * 0.5;
  out->pair_count = hh / 2;
}

}  // namespace ljg
