// ljg_api.hip -- the C ABI (include/ljgpu.h): context lifetime, uploads/downloads in upload order, the
// rebuild policy and the step loop.  All HIP errors are caught here and turned into status codes.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "ljg_internal.h"

struct ljg_ctx {
  ljg::Ctx c;
};

namespace ljg {

static ljg_status fail(Ctx &c, ljg_status st, const std::string &msg) {
  c.err = msg;
  return st;
}

#define LJG_API_BEGIN(ctxp)                  \
  if (!(ctxp)) return LJG_ERR_ARG;           \
  Ctx &c = (ctxp)->c;                        \
  try {                                      \
    LJG_HIP(hipSetDevice(c.device));
#define LJG_API_END                                                                                  \
  }                                                                                                  \
  catch (const HipError &e) {                                                                        \
    char buf[512];                                                                                   \
    std::snprintf(buf, sizeof buf, "%s failed: %s (line %d)", e.what, hipGetErrorString(e.err), e.line); \
    return fail(c, e.line == -2 ? LJG_ERR_UNSUPPORTED : e.line < 0 ? LJG_ERR_CAPACITY : LJG_ERR_HIP, buf); \
  }                                                                                                  \
  catch (const std::bad_alloc &) {                                                                   \
    return fail(c, LJG_ERR_HIP, "host allocation failed");                                          \
  }

__global__ void k_init_upload(int64_t n, StateView s, const int64_t *__restrict__ id_in, int32_t have_id) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  s.fx[i] = 0.; s.fy[i] = 0.; s.fz[i] = 0.;
  s.ofx[i] = 0.; s.ofy[i] = 0.; s.ofz[i] = 0.;
  s.rx[i] = s.x[i]; s.ry[i] = s.y[i]; s.rz[i] = s.z[i];
  s.orig[i] = (int32_t)i;
  s.id[i] = have_id ? id_in[i] : i;
}

// Positions, velocities and forces of the particles with upload indices [first, first + count), array by
// array (x, y, z, vx, vy, vz, fx, fy, fz), count values each.
__global__ void k_gather_range(int64_t nslots, const int32_t *__restrict__ orig, int64_t first, int64_t count,
                               StateView s, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  const int64_t k = (int64_t)orig[i] - first;
  if (k < 0 || k >= count) return;
  const double *in[9] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.fx, s.fy, s.fz};
#pragma unroll
  for (int a = 0; a < 9; ++a) out[a * count + k] = in[a][i];
}

// out[s] = in[orig[s]] (upload order -> storage order).
__global__ void k_from_upload_order(int64_t nslots, const int32_t *__restrict__ orig, const double *__restrict__ in,
                                    double *__restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const int32_t o = orig[s];
  if (o >= 0) out[s] = in[o];
}

__global__ void k_zero_i64(int64_t n, int64_t *p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_zero_i32(int64_t n, int32_t *p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

static bool is_n3_kernel(int k) { return k == LJG_KERNEL_CLUSTER_FP64_N3 || k == LJG_KERNEL_CLUSTER_MIXED_N3; }

static bool is_cluster_kernel(int k) {
  return k == LJG_KERNEL_CLUSTER_FP64 || k == LJG_KERNEL_CLUSTER_MIXED || k == LJG_KERNEL_CLUSTER_JROWS_MIXED ||
         is_n3_kernel(k);
}

static void rebuild_for(Ctx &c, int kernel) {
  if (is_cluster_kernel(kernel)) {  // the tower histogram wraps the positions
    c.want_jrows = kernel == LJG_KERNEL_CLUSTER_JROWS_MIXED;
    c.want_n3 = is_n3_kernel(kernel);
    vcl_build(c);
  } else {
    wrap_positions(c);
    lc_bin(c);
    if (kernel == LJG_KERNEL_VERLET_MIXED) verlet_build(c);
  }
  c.steps_since_rebuild = 0;  // the rebuild's permutation set the reference positions
  std::fill(c.box_since.begin(), c.box_since.end(), 0);
  c.box_pending.clear();
}

static bool layout_matches(const Ctx &c, int kernel);

// Rebuild of the ensemble boxes `mark` lists (host, per box), in place in their regions; the other boxes
// keep their layout and lists.  Every box marked: the whole build.
// The marked boxes' slots, towers and clusters as segments and their list, the clusters of the other
// boxes as segments (runs of consecutive unmarked boxes), and the mark itself, uploaded on the stream for
// a partial build and the force launches of the two sets.  False when no box is marked.
static bool prepare_segments(Ctx &c, const std::vector<int32_t> &mark) {
  std::vector<int32_t> act;
  for (int32_t b = 0; b < (int32_t)mark.size(); ++b)
    if (mark[b]) act.push_back(b);
  if (act.empty()) return false;
  const int32_t n = (int32_t)act.size(), M = c.M, nb = (int32_t)mark.size();
  const std::vector<int64_t> &cb = c.box_cbase_h, &tb = c.box_tbase_h;
  int32_t runs = 0;
  for (int32_t b = 0; b < nb; ++b) runs += !mark[b] && (b == 0 || mark[b - 1]);
  // [prefix n+1, base n] for slots, towers, clusters; the boxes; [prefix runs+1, base runs] of the others
  std::vector<int64_t> &h = c.seg_host;
  h.assign(7 * (size_t)n + 3 + 2 * (size_t)runs + 1, 0);
  int64_t *sp = h.data(), *sb = sp + n + 1, *tp = sb + n, *tbs = tp + n + 1, *cp = tbs + n, *cbs = cp + n + 1;
  int64_t *bx = cbs + n, *op = bx + n, *ob = op + runs + 1;
  // longest tower expected: twice a box's mean tower plus 32, at most its particles (a longer tower is
  // reported by the build, which then reads the longest tower back and runs again)
  c.seg_max_count = 0;
  for (int32_t j = 0; j < n; ++j) {
    const int32_t b = act[j];
    const int64_t nt = tb[b + 1] - tb[b], mean = (c.box_count[b] + nt - 1) / nt;
    const int64_t guess = 2 * mean + 32 < c.box_count[b] ? 2 * mean + 32 : c.box_count[b];
    c.seg_max_count = guess > c.seg_max_count ? guess : c.seg_max_count;
    sb[j] = cb[b] * M;
    sp[j + 1] = sp[j] + (cb[b + 1] - cb[b]) * M;
    tbs[j] = tb[b];
    tp[j + 1] = tp[j] + (tb[b + 1] - tb[b]);
    cbs[j] = cb[b];
    cp[j + 1] = cp[j] + (cb[b + 1] - cb[b]);
    bx[j] = b;
  }
  for (int32_t b = 0, r = 0; b < nb;) {
    if (mark[b]) {
      ++b;
      continue;
    }
    int32_t e = b;
    while (e < nb && !mark[e]) ++e;
    ob[r] = cb[b];
    op[r + 1] = op[r] + (cb[e] - cb[b]);
    ++r;
    b = e;
  }
  // one copy from pinned memory: the segments, then the mark (int32 per box)
  const size_t bytes = h.size() * sizeof(int64_t) + (size_t)nb * sizeof(int32_t);
  if (c.pin_seg_cap < bytes) {
    if (c.pin_seg) LJG_HIP(hipHostFree(c.pin_seg));
    c.pin_seg = nullptr;
    c.pin_seg_cap = 0;
    LJG_HIP(hipHostMalloc((void **)&c.pin_seg, bytes + bytes / 2, hipHostMallocDefault));
    c.pin_seg_cap = bytes + bytes / 2;
  }
  std::memcpy(c.pin_seg, h.data(), h.size() * sizeof(int64_t));
  std::memcpy(c.pin_seg + h.size() * sizeof(int64_t), mark.data(), (size_t)nb * sizeof(int32_t));
  c.seg_buf.ensure((bytes + sizeof(int64_t) - 1) / sizeof(int64_t));
  const int64_t *d = c.seg_buf.p;
  LJG_HIP(hipMemcpyAsync(c.seg_buf.p, c.pin_seg, bytes, hipMemcpyHostToDevice, c.stream));
  c.seg_active = reinterpret_cast<const int32_t *>(d + h.size());
  auto dev = [&](const int64_t *p) { return d + (p - h.data()); };
  c.seg_slots = Segs{dev(sp), dev(sb), n};
  c.seg_towers = Segs{dev(tp), dev(tbs), n};
  c.seg_clusters = Segs{dev(cp), dev(cbs), n};
  c.seg_nslots = sp[n];
  c.seg_ntowers = tp[n];
  c.seg_nclusters = cp[n];
  c.seg_boxes = dev(bx);
  c.seg_nboxes = n;
  c.seg_others = Segs{dev(op), dev(ob), runs};
  c.seg_nothers = op[runs];
  return true;
}

// Rebuild of the ensemble boxes `mark` lists (host, per box), in place in their regions; the other boxes
// keep their layout and lists.  Every box marked: the whole build.  segments_ready: prepare_segments ran
// for this mark already.
static void rebuild_boxes(Ctx &c, int kernel, const std::vector<int32_t> &mark, bool segments_ready = false) {
  bool all = true;
  for (int32_t m : mark) all = all && m != 0;
  if (all || !c.regions_valid || !layout_matches(c, kernel)) {
    rebuild_for(c, kernel);
    return;
  }
  if (!segments_ready && !prepare_segments(c, mark)) return;
  c.want_jrows = false;
  c.want_n3 = false;
  c.build_partial = true;
  try {
    vcl_build(c);
  } catch (...) {
    c.build_partial = false;
    throw;
  }
  c.build_partial = false;
  for (size_t b = 0; b < mark.size(); ++b)
    if (mark[b]) c.box_since[b] = 0;
}

// The layout a kernel needs: the whole build when the layout is another, else the boxes an edit added.
// Returns whether anything was built.
static bool prepare_layout(Ctx &c, int kernel) {
  if (!layout_matches(c, kernel)) {
    rebuild_for(c, kernel);
    return true;
  }
  bool pending = false;
  for (int32_t p : c.box_pending) pending = pending || p != 0;
  if (!pending) return false;
  const std::vector<int32_t> mark = c.box_pending;
  rebuild_boxes(c, kernel, mark);
  c.box_pending.clear();
  return true;
}

static bool layout_matches(const Ctx &c, int kernel) {
  if (is_cluster_kernel(kernel)) {
    const bool entries = kernel != LJG_KERNEL_CLUSTER_JROWS_MIXED && !is_n3_kernel(kernel);
    return c.layout == Layout::Clusters && c.built_M == c.M &&
           c.jrows_valid == (kernel == LJG_KERNEL_CLUSTER_JROWS_MIXED) && c.lists_n3 == is_n3_kernel(kernel) &&
           (!entries || c.xrows_valid == !c.prm.whole_cluster_rows);
  }
  if (kernel == LJG_KERNEL_VERLET_MIXED) return c.layout == Layout::Verlet;
  return c.layout == Layout::Cells;
}

// The fixed-point kernels rely on minimum-image differences: the Verlet-list kernel also on 27 distinct
// neighbour cells (box >= 3 ilen), the j-row cluster kernel on a single image within ilen (box > 2 ilen).
static const char *kernel_unsupported(const Ctx &c, int kernel) {
  if (c.nboxes > 0 && kernel != LJG_KERNEL_CLUSTER_MIXED)
    return "an ensemble of boxes runs the FP32 cluster kernel only (without Newton-3: its boxes stay bit-identical "
           "to each box alone)";
  const double ilen = c.prm.cutoff + c.prm.skin;
  for (int d = 0; d < 3; ++d) {
    const double L = c.prm.box_max[d] - c.prm.box_min[d];
    if (kernel == LJG_KERNEL_VERLET_MIXED && !(L >= 3.0 * ilen)) return "box shorter than 3 interaction lengths";
    if (kernel == LJG_KERNEL_CLUSTER_JROWS_MIXED && !(L > 2.0 * ilen))
      return "box not longer than 2 interaction lengths";
  }
  return nullptr;
}

static int64_t partial_items(const Ctx &c, int /*kernel*/) { return c.part_items; }

static void run_force(Ctx &c, int kernel, bool globals) {
  if (kernel == LJG_KERNEL_LC_FP64) lc_force(c, globals);
  else if (kernel == LJG_KERNEL_VERLET_MIXED) verlet_force(c, globals);
  else if (kernel == LJG_KERNEL_CLUSTER_JROWS_MIXED) cluster_jrows_force(c, globals);
  else
    cluster_force(c, kernel == LJG_KERNEL_CLUSTER_MIXED || kernel == LJG_KERNEL_CLUSTER_MIXED_N3, globals, nullptr,
                  nullptr, is_n3_kernel(kernel));
}

static int64_t candidate_evals(const Ctx &c, int kernel) {
  if (kernel == LJG_KERNEL_CLUSTER_JROWS_MIXED) return c.nentries * (int64_t)c.M;
  if (is_cluster_kernel(kernel)) return c.nentries * (int64_t)c.M * c.M;
  if (kernel == LJG_KERNEL_VERLET_MIXED) return c.nentries;
  return -1;
}

}  // namespace ljg

using namespace ljg;

extern "C" {

ljg_status ljg_create(const ljg_params *p, ljg_ctx **out) {
  if (!p || !out) return LJG_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return LJG_ERR_NODEVICE;
  if (p->device < 0 || p->device >= ndev) return LJG_ERR_ARG;
  const double ilen = p->cutoff + p->skin;
  if (!(p->cutoff > 0.) || !(p->skin >= 0.)) return LJG_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (!(p->box_max[d] - p->box_min[d] >= ilen)) return LJG_ERR_ARG;  // LogicHandler.h:967-983
  if (p->cluster_size != 1 && p->cluster_size != 2 && p->cluster_size != 4 && p->cluster_size != 8)
    return LJG_ERR_UNSUPPORTED;
  if (p->rebuild_frequency < 1) return LJG_ERR_ARG;
  if (p->tile_groups < 0 || p->tile_groups > 64) return LJG_ERR_ARG;
  if (p->build_groups < 0 || p->build_groups > 64) return LJG_ERR_ARG;
  if (p->list_mode < 0 || p->list_mode > 3) return LJG_ERR_ARG;
  if (p->shared_rebuilds < 0 || p->shared_rebuilds > 1) return LJG_ERR_ARG;
  if (p->graphs < 0 || p->graphs > 1) return LJG_ERR_ARG;
  if (p->row_headroom < -1 || p->row_headroom > 400) return LJG_ERR_ARG;
  ljg_ctx *h = new (std::nothrow) ljg_ctx;
  if (!h) return LJG_ERR_HIP;
  Ctx &c = h->c;
  c.prm = *p;
  c.device = p->device;
  c.M = p->cluster_size;
  try {
    LJG_HIP(hipSetDevice(c.device));
    LJG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    LJG_HIP(hipEventCreate(&c.ev0));
    LJG_HIP(hipEventCreate(&c.ev1));
    // coherent, device-mapped: the drift kernel writes the displacement stamp here directly
    LJG_HIP(hipHostMalloc((void **)&c.host_flag, 4 * sizeof(int32_t), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(c.host_flag, 0, 4 * sizeof(int32_t));
    LJG_HIP(hipHostGetDevicePointer((void **)&c.host_flag_dev, c.host_flag, 0));
    LJG_HIP(hipHostMalloc((void **)&c.pin_mark, (1 + kMarkWords) * sizeof(unsigned long long),
                          hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(c.pin_mark, 0, (1 + kMarkWords) * sizeof(unsigned long long));
    LJG_HIP(hipHostGetDevicePointer((void **)&c.pin_mark_dev, c.pin_mark, 0));
    c.flag.ensure(4);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, c.stream, (int64_t)4, c.flag.p);
    LJG_HIP(hipGetLastError());
  } catch (const HipError &e) {
    delete h;
    return LJG_ERR_HIP;
  }
  const double eps1 = 1., sig1 = 1., mass1 = 1.;
  const ljg_status st = ljg_set_types(h, 1, &eps1, &sig1, &mass1);
  if (st != LJG_OK) {
    ljg_destroy(h);
    return st;
  }
  *out = h;
  return LJG_OK;
}

ljg_status ljg_destroy(ljg_ctx *h) {
  if (!h) return LJG_ERR_ARG;
  Ctx &c = h->c;
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  if (c.side) (void)hipStreamSynchronize(c.side);
  for (auto &s : c.st) s.release();
  c.cell_of.release(); c.cell_count.release(); c.cell_start.release(); c.rank.release(); c.perm.release();
  c.cell_of_particle.release(); c.tower_of.release(); c.tower_count.release(); c.tower_start.release();
  c.tower_items.release(); c.tower_z.release(); c.tower_id.release(); c.ctow_count.release(); c.ctow_start.release(); c.cluster_tower.release();
  c.bbox.release(); c.c0.release(); c.list_count.release(); c.list_stats.release(); c.list.release();
  c.pos32.release(); c.pos32_next.release(); c.posd.release(); c.posd_next.release(); c.pos32_ext.release(); c.posfx.release(); c.jcl.release(); c.jlist.release(); c.jcount.release(); c.xrow.release(); c.xcount.release(); c.xrow_tmp.release(); c.xcount_tmp.release(); c.tstats.release(); c.posfx_next.release(); c.hits_slots.release(); c.vcount.release(); c.vlist.release(); c.tab.release(); c.part_e.release(); c.part_v.release(); c.part_hits.release();
  c.red_out.release(); c.red_hits.release(); c.flag.release(); c.scan_tmp.release();
  c.box_tg.release(); c.box_wrap.release(); c.box_of.release(); c.box_tbase.release(); c.tower_box.release();
  c.box_cbase.release(); c.box_pbase.release(); c.cluster_box.release();
  c.box_flag.release(); c.list_tmp.release(); c.seg_buf.release(); c.cell_sorted.release();
  c.pos32_alt.release(); c.pos32n_alt.release(); c.posd_alt.release(); c.posdn_alt.release(); c.bbox_alt.release(); c.c0_alt.release(); c.lcount_alt.release();
  c.ctower_alt.release();
  c.cl_glob.release(); c.box_glob.release();
  if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
  if (c.host_flag) (void)hipHostFree(c.host_flag);
  if (c.pin_mark) (void)hipHostFree(c.pin_mark);
  if (c.pin_seg) (void)hipHostFree(c.pin_seg);
  if (c.pin_dl) (void)hipHostFree(c.pin_dl);
  if (c.pin_flags) (void)hipHostFree(c.pin_flags);
  for (auto e : c.event_pool) (void)hipEventDestroy(e);
  if (c.ev0) (void)hipEventDestroy(c.ev0);
  if (c.ev1) (void)hipEventDestroy(c.ev1);
  if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
  if (c.ev_join) (void)hipEventDestroy(c.ev_join);
  if (c.side) (void)hipStreamDestroy(c.side);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
  return LJG_OK;
}

const char *ljg_last_error(const ljg_ctx *h) { return h ? h->c.err.c_str() : "null context"; }

ljg_status ljg_device_name(const ljg_ctx *h, char *buf, int32_t buflen) {
  if (!h || !buf || buflen <= 0) return LJG_ERR_ARG;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, h->c.device) != hipSuccess) return LJG_ERR_HIP;
  std::snprintf(buf, (size_t)buflen, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return LJG_OK;
}

static void mix_types(Ctx &c, int32_t ntypes, const double *eps, const double *sig, const double *mass);

ljg_status ljg_memory_info(const ljg_ctx *h, uint64_t *free_bytes, uint64_t *total_bytes) {
  if (!h || !free_bytes || !total_bytes) return LJG_ERR_ARG;
  if (hipSetDevice(h->c.device) != hipSuccess) return LJG_ERR_HIP;
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return LJG_ERR_HIP;
  *free_bytes = f;
  *total_bytes = t;
  return LJG_OK;
}

ljg_status ljg_set_types(ljg_ctx *h, int32_t ntypes, const double *eps, const double *sig, const double *mass) {
  if (!h || ntypes < 1 || ntypes > kMaxTypes || !eps || !sig) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  mix_types(c, ntypes, eps, sig, mass);
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_set_layout(ljg_ctx *h, int32_t cluster_size, int32_t tile_groups, int32_t build_groups) {
  if (!h) return LJG_ERR_ARG;
  if (cluster_size != 1 && cluster_size != 2 && cluster_size != 4 && cluster_size != 8) return LJG_ERR_UNSUPPORTED;
  if (tile_groups < 0 || tile_groups > 64 || build_groups < 0 || build_groups > 64) return LJG_ERR_ARG;
  Ctx &c = h->c;
  if (c.nboxes > 0 && cluster_size != c.M) return fail(c, LJG_ERR_UNSUPPORTED, "an ensemble keeps its cluster size");
  c.M = cluster_size;
  c.prm.cluster_size = cluster_size;
  c.prm.tile_groups = tile_groups;
  c.prm.build_groups = build_groups;
  return LJG_OK;
}

ljg_status ljg_set_shift(ljg_ctx *h, int32_t apply_shift) {
  LJG_API_BEGIN(h)
  c.prm.apply_shift = apply_shift != 0;
  const LJParams old = c.lj;
  mix_types(c, old.ntypes, old.eps, old.sig, old.mass);
  return LJG_OK;
  LJG_API_END
}

static void mix_types(Ctx &c, int32_t ntypes, const double *eps, const double *sig, const double *mass) {
  LJParams &lj = c.lj;
  for (int a = 0; a < ntypes; ++a) {
    lj.eps[a] = eps[a];
    lj.sig[a] = sig[a];
  }
  lj.ntypes = ntypes;
  const double rc2 = c.prm.cutoff * c.prm.cutoff;
  lj.rc2 = rc2;
  mix_table(ntypes, eps, sig, rc2, c.prm.apply_shift != 0, lj.eps24, lj.sig2, lj.shift6);
  for (int t = 0; t < kMaxTypes; ++t) {
    lj.mass[t] = (mass && t < ntypes) ? mass[t] : 1.;
    lj.inv2m[t] = 1. / (2. * lj.mass[t]);
  }
  std::vector<double> host(3 * ntypes * ntypes);
  for (int k = 0; k < ntypes * ntypes; ++k) {
    host[3 * k] = lj.eps24[k];
    host[3 * k + 1] = lj.sig2[k];
    host[3 * k + 2] = lj.shift6[k];
  }
  c.tab.ensure(host.size());
  LJG_HIP(hipMemcpyAsync(c.tab.p, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  LJG_HIP(hipStreamSynchronize(c.stream));
  c.types_set = true;
}

// The periodic box the kernels work in; the particles are dropped.
static void reset_box(Ctx &c, const double box_min[3], const double box_max[3]) {
  LJG_HIP(hipStreamSynchronize(c.stream));
  for (int d = 0; d < 3; ++d) {
    c.prm.box_min[d] = box_min[d];
    c.prm.box_max[d] = box_max[d];
  }
  c.nboxes = 0;
  c.box_glob_valid = false;
  c.n = 0;
  c.nslots = 0;
  c.n_owned = -1;
  c.layout = Layout::None;
  c.have_forces = false;
  c.steps_since_rebuild = 0;
}

ljg_status ljg_set_open_box(ljg_ctx *h, const double box_min[3], const double box_max[3]) {
  if (!h || !box_min || !box_max) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  const double ilen = c.prm.cutoff + c.prm.skin;
  for (int d = 0; d < 3; ++d)
    if (!(box_max[d] > box_min[d])) return fail(c, LJG_ERR_ARG, "empty box");
  // the box and its halo region [min - ilen, max + ilen], then ilen of empty space before the periodic
  // image of the low halo edge: two particles of the region are at least ilen apart across the seam
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = box_min[d] - ilen;
    hi[d] = box_max[d] + 2.0 * ilen;
  }
  reset_box(c, lo, hi);
  c.open = true;
  for (int d = 0; d < 3; ++d) {
    c.open_min[d] = box_min[d];
    c.open_max[d] = box_max[d];
  }
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_set_owned(ljg_ctx *h, int64_t n_owned) {
  if (!h) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.nboxes > 0) return fail(c, LJG_ERR_UNSUPPORTED, "ownership on an ensemble of boxes");
  if (n_owned < 0 || n_owned > c.n) return fail(c, LJG_ERR_ARG, "owned count outside [0, uploaded particles]");
  c.n_owned = n_owned;
  c.box_glob_valid = false;
  return LJG_OK;
  LJG_API_END
}

// An open box takes particles in its halo region only (host arrays, checked before any copy).
static const char *outside_open_region(const Ctx &c, int64_t n, const double *x, const double *y,
                                       const double *z) {
  if (!c.open) return nullptr;
  const double ilen = c.prm.cutoff + c.prm.skin;
  const double *p[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) {
    const double lo = c.open_min[d] - ilen, hi = c.open_max[d] + ilen;
    for (int64_t i = 0; i < n; ++i)
      if (!(p[d][i] >= lo && p[d][i] <= hi)) return "particle outside the open box and its halo region";
  }
  return nullptr;
}

ljg_status ljg_set_box(ljg_ctx *h, const double box_min[3], const double box_max[3]) {
  if (!h || !box_min || !box_max) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  const double ilen = c.prm.cutoff + c.prm.skin;
  for (int d = 0; d < 3; ++d)
    if (!(box_max[d] - box_min[d] >= ilen)) return fail(c, LJG_ERR_ARG, "box shorter than rc + skin");
  c.open = false;
  reset_box(c, box_min, box_max);
  return LJG_OK;
  LJG_API_END
}

// Box list checks shared by ljg_set_boxes and ljg_edit_boxes; returns the reason or nullptr.
static const char *check_boxes(const Ctx &c, int32_t nboxes, const double *box_min, const double *box_max,
                               const int64_t *count, int64_t *total) {
  const double ilen = c.prm.cutoff + c.prm.skin;
  int64_t t = 0;
  for (int32_t b = 0; b < nboxes; ++b) {
    if (count[b] < 1) return "a box without particles";
    for (int d = 0; d < 3; ++d)
      if (!(box_max[3 * b + d] - box_min[3 * b + d] >= ilen)) return "box shorter than rc + skin";
    t += count[b];
  }
  if (t > (int64_t)INT32_MAX / 2) return "too many particles";
  *total = t;
  return nullptr;
}

// Host and device tables of an ensemble (bounds, counts, box of each particle, wrap arguments).
static void set_box_tables(Ctx &c, int32_t nboxes, const double *box_min, const double *box_max,
                           const int64_t *count, int64_t total) {
  std::vector<double> lo(box_min, box_min + 3 * (size_t)nboxes), hi(box_max, box_max + 3 * (size_t)nboxes);
  std::vector<int64_t> cnt(count, count + nboxes), pstart((size_t)nboxes + 1, 0);
  std::vector<WrapArgs> wr((size_t)nboxes);
  for (int32_t b = 0; b < nboxes; ++b) {
    pstart[b + 1] = pstart[b] + cnt[b];
    wr[b] = wrap_args(&lo[3 * b], &hi[3 * b]);
  }
  c.box_of.ensure((size_t)total);
  c.box_wrap.ensure(wr.size());
  c.box_pbase.ensure(pstart.size());
  LJG_HIP(hipMemcpyAsync(c.box_wrap.p, wr.data(), wr.size() * sizeof(WrapArgs), hipMemcpyHostToDevice, c.stream));
  LJG_HIP(hipMemcpyAsync(c.box_pbase.p, pstart.data(), pstart.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                         c.stream));
  fill_owner(c, c.box_of.p, total, c.box_pbase.p, nboxes);  // box of each particle, upload order
  c.box_lo = std::move(lo);
  c.box_hi = std::move(hi);
  c.box_count = std::move(cnt);
  // the single-box parameters keep the first box (the fixed-point and cell geometry never run here)
  for (int d = 0; d < 3; ++d) {
    c.prm.box_min[d] = c.box_lo[d];
    c.prm.box_max[d] = c.box_hi[d];
  }
  c.nboxes = nboxes;
  c.box_geom_M = 0;
  c.box_glob_valid = false;
  c.regions_valid = false;
  c.build_partial = false;
  c.box_since.assign((size_t)nboxes, 0);
  c.box_pending.clear();
  c.box_flag.ensure((size_t)nboxes + 1);  // (+1: read back as whole 8-byte words)
  fill_i32(c, c.box_flag.p, nboxes + 1, 0);
  LJG_HIP(hipStreamSynchronize(c.stream));  // the host tables go out of scope
}

ljg_status ljg_set_boxes(ljg_ctx *h, int32_t nboxes, const double *box_min, const double *box_max,
                         const int64_t *count) {
  if (!h || nboxes < 1 || !box_min || !box_max || !count) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  int64_t total = 0;
  if (const char *why = check_boxes(c, nboxes, box_min, box_max, count, &total)) return fail(c, LJG_ERR_ARG, why);
  LJG_HIP(hipStreamSynchronize(c.stream));
  set_box_tables(c, nboxes, box_min, box_max, count, total);
  c.open = false;  // the boxes are periodic
  c.n_owned = -1;
  c.n = 0;
  c.nslots = 0;
  c.layout = Layout::None;
  c.have_forces = false;
  c.steps_since_rebuild = 0;
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_edit_boxes(ljg_ctx *h, int32_t nkeep, const int32_t *keep, int32_t nadd, const double *box_min,
                          const double *box_max, const int64_t *count, const double *x, const double *y,
                          const double *z, const double *vx, const double *vy, const double *vz, const int32_t *type,
                          const int64_t *id) {
  if (!h || nkeep < 0 || nadd < 0 || nkeep + nadd < 1 || (nkeep > 0 && !keep)) return LJG_ERR_ARG;
  if (nadd > 0 && (!box_min || !box_max || !count || !x || !y || !z)) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (nkeep > 0 && (c.nboxes == 0 || c.n == 0)) return fail(c, LJG_ERR_STATE, "no ensemble to keep boxes from");
  if (c.open) return fail(c, LJG_ERR_UNSUPPORTED, "an open box is one domain, not an ensemble");
  const int32_t old_nb = c.nboxes;
  std::vector<int64_t> old_pstart((size_t)old_nb + 1, 0), new_base((size_t)old_nb, -1);
  for (int32_t b = 0; b < old_nb; ++b) old_pstart[b + 1] = old_pstart[b] + c.box_count[b];
  const int32_t nb = nkeep + nadd;
  std::vector<double> lo(3 * (size_t)nb), hi(3 * (size_t)nb);
  std::vector<int64_t> cnt((size_t)nb);
  int64_t kept = 0;
  for (int32_t k = 0; k < nkeep; ++k) {
    const int32_t b = keep[k];
    if (b < 0 || b >= old_nb || new_base[b] >= 0) return fail(c, LJG_ERR_ARG, "keep lists a box that does not exist or twice");
    new_base[b] = kept;
    for (int d = 0; d < 3; ++d) {
      lo[3 * k + d] = c.box_lo[3 * b + d];
      hi[3 * k + d] = c.box_hi[3 * b + d];
    }
    cnt[k] = c.box_count[b];
    kept += cnt[k];
  }
  int64_t added = 0;
  for (int32_t a = 0; a < nadd; ++a) {
    for (int d = 0; d < 3; ++d) {
      lo[3 * (nkeep + a) + d] = box_min[3 * a + d];
      hi[3 * (nkeep + a) + d] = box_max[3 * a + d];
    }
    cnt[nkeep + a] = count[a];
    added += count[a];
  }
  if (type) {
    for (int64_t i = 0; i < added; ++i)
      if (type[i] < 0 || type[i] >= c.lj.ntypes) return fail(c, LJG_ERR_ARG, "particle type not set");
  }
  int64_t total = 0;
  if (const char *why = check_boxes(c, nb, lo.data(), hi.data(), cnt.data(), &total)) return fail(c, LJG_ERR_ARG, why);
  if (c.nboxes > 0 && c.regions_valid && c.layout == Layout::Clusters && !c.jrows_valid && c.built_M == c.M) {
    // the running boxes keep their layouts, lists and rebuild schedules; only the added boxes are built
    // (before the next force), so every box runs as it would alone
    const std::vector<int32_t> src(keep, keep + nkeep);
    const std::vector<int64_t> since = c.box_since;
    const std::vector<int32_t> pending = c.box_pending;
    set_box_tables(c, nb, lo.data(), hi.data(), cnt.data(), total);
    ensemble_edit_regions(c, src, old_pstart, kept, added, x, y, z, vx, vy, vz, type, id);
    c.box_pending.assign((size_t)nb, 1);
    for (int32_t k = 0; k < nkeep; ++k) {
      c.box_since[k] = since[src[k]];
      c.box_pending[k] = pending.empty() ? 0 : pending[src[k]];
    }
    c.n = total;
    c.have_forces = false;
    return LJG_OK;
  }
  ensemble_edit(c, old_pstart, new_base, kept, added, x, y, z, vx, vy, vz, type, id);
  set_box_tables(c, nb, lo.data(), hi.data(), cnt.data(), total);
  c.n = total;
  c.nslots = total;
  c.layout = Layout::Upload;
  c.regions_valid = false;
  c.have_forces = false;
  c.steps_since_rebuild = 0;
  return LJG_OK;
  LJG_API_END
}

// Upload indices [first, first + count): one pass gathers the nine arrays (null outputs are skipped on the
// host), one copy brings them back through pinned memory.
static void gather_range(Ctx &c, int64_t first, int64_t count, double *x, double *y, double *z, double *vx,
                         double *vy, double *vz, double *fx, double *fy, double *fz) {
  State &s = c.st[c.cur];
  double *outs[9] = {x, y, z, vx, vy, vz, fx, fy, fz};
  const size_t bytes = 9 * (size_t)count * sizeof(double);
  c.part_v.ensure(9 * (size_t)count);
  hipLaunchKernelGGL(k_gather_range, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots, s.orig.p,
                     first, count, view(s), c.part_v.p);
  LJG_HIP(hipGetLastError());
  if (c.pin_dl_cap < bytes) {
    if (c.pin_dl) LJG_HIP(hipHostFree(c.pin_dl));
    c.pin_dl = nullptr;
    c.pin_dl_cap = 0;
    LJG_HIP(hipHostMalloc((void **)&c.pin_dl, bytes + bytes / 2, hipHostMallocDefault));
    c.pin_dl_cap = bytes + bytes / 2;
  }
  LJG_HIP(hipMemcpyAsync(c.pin_dl, c.part_v.p, bytes, hipMemcpyDeviceToHost, c.stream));
  stream_wait(c);
  for (int k = 0; k < 9; ++k)
    if (outs[k]) std::memcpy(outs[k], c.pin_dl + (size_t)k * count, (size_t)count * sizeof(double));
}

ljg_status ljg_download_range(ljg_ctx *h, int64_t first, int64_t count, double *x, double *y, double *z, double *vx,
                              double *vy, double *vz, double *fx, double *fy, double *fz) {
  if (!h || first < 0 || count < 0) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  if (first + count > c.n) return fail(c, LJG_ERR_ARG, "range beyond the uploaded particles");
  if (count == 0) return LJG_OK;
  gather_range(c, first, count, x, y, z, vx, vy, vz, fx, fy, fz);
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_get_box_globals(ljg_ctx *h, int32_t nboxes, double *potential_energy, double *virial,
                               int64_t *pair_count) {
  if (!h) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.nboxes == 0 || nboxes != c.nboxes) return fail(c, LJG_ERR_ARG, "not an ensemble of that many boxes");
  if (!c.box_glob_valid) return fail(c, LJG_ERR_STATE, "no force evaluation with globals since the last change");
  std::vector<double> g(3 * (size_t)nboxes);
  LJG_HIP(hipMemcpyAsync(g.data(), c.box_glob.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  LJG_HIP(hipStreamSynchronize(c.stream));
  for (int32_t b = 0; b < nboxes; ++b) {
    if (potential_energy) potential_energy[b] = g[3 * b];
    if (virial) virial[b] = g[3 * b + 1];
    // half the box's pair hits, rounded down as a box alone reports them (a pair at the cutoff can be
    // counted from one side only: the FP32 separations of the two directions round apart)
    if (pair_count) pair_count[b] = (int64_t)std::floor(g[3 * b + 2]);
  }
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_upload(ljg_ctx *h, int64_t n, const double *x, const double *y, const double *z, const double *vx,
                      const double *vy, const double *vz, const int32_t *type, const int64_t *id) {
  if (!h || n <= 0 || n > (int64_t)INT32_MAX / 2 || !x || !y || !z) return LJG_ERR_ARG;
  if (type) {
    for (int64_t i = 0; i < n; ++i)
      if (type[i] < 0 || type[i] >= h->c.lj.ntypes) return LJG_ERR_ARG;
  }
  LJG_API_BEGIN(h)
  if (c.nboxes > 0) {
    int64_t total = 0;
    for (int64_t k : c.box_count) total += k;
    if (n != total) return fail(c, LJG_ERR_ARG, "particle count differs from the ensemble's boxes");
  }
  if (const char *why = outside_open_region(c, n, x, y, z)) return fail(c, LJG_ERR_ARG, why);
  c.box_glob_valid = false;
  c.n_owned = -1;
  c.cur = 0;
  c.st[0].ensure((size_t)n);
  c.n = n;
  c.nslots = n;
  State &s = c.st[0];
  const size_t bytes = (size_t)n * sizeof(double);
  LJG_HIP(hipMemcpyAsync(s.x.p, x, bytes, hipMemcpyHostToDevice, c.stream));
  LJG_HIP(hipMemcpyAsync(s.y.p, y, bytes, hipMemcpyHostToDevice, c.stream));
  LJG_HIP(hipMemcpyAsync(s.z.p, z, bytes, hipMemcpyHostToDevice, c.stream));
  if (vx && vy && vz) {
    LJG_HIP(hipMemcpyAsync(s.vx.p, vx, bytes, hipMemcpyHostToDevice, c.stream));
    LJG_HIP(hipMemcpyAsync(s.vy.p, vy, bytes, hipMemcpyHostToDevice, c.stream));
    LJG_HIP(hipMemcpyAsync(s.vz.p, vz, bytes, hipMemcpyHostToDevice, c.stream));
  } else {
    LJG_HIP(hipMemsetAsync(s.vx.p, 0, bytes, c.stream));
    LJG_HIP(hipMemsetAsync(s.vy.p, 0, bytes, c.stream));
    LJG_HIP(hipMemsetAsync(s.vz.p, 0, bytes, c.stream));
  }
  if (type) LJG_HIP(hipMemcpyAsync(s.type.p, type, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  else LJG_HIP(hipMemsetAsync(s.type.p, 0, (size_t)n * sizeof(int32_t), c.stream));
  int64_t *did = nullptr;
  if (id) {
    c.scan_tmp.ensure((size_t)n);
    did = c.scan_tmp.p;
    LJG_HIP(hipMemcpyAsync(did, id, (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
  }
  hipLaunchKernelGGL(k_init_upload, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, view(s), did,
                     (int32_t)(id != nullptr));
  LJG_HIP(hipGetLastError());
  if (c.nboxes > 0) ensemble_wrap(c);
  else wrap_positions(c);
  stream_wait(c);  // the caller's arrays are free to change once this returns
  c.layout = Layout::Upload;
  c.regions_valid = false;
  c.have_forces = false;
  c.steps_since_rebuild = 0;
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_set_forces(ljg_ctx *h, const double *fx, const double *fy, const double *fz) {
  if (!h || !fx || !fy || !fz) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  State &s = c.st[c.cur];
  c.part_e.ensure((size_t)c.n);
  const double *src[3] = {fx, fy, fz};
  double *dst[3] = {s.fx.p, s.fy.p, s.fz.p};
  for (int d = 0; d < 3; ++d) {
    LJG_HIP(hipMemcpyAsync(c.part_e.p, src[d], (size_t)c.n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_from_upload_order, dim3(blocks_for(c.nslots, 256)), dim3(256), 0, c.stream, c.nslots,
                       s.orig.p, c.part_e.p, dst[d]);
    LJG_HIP(hipGetLastError());
  }
  LJG_HIP(hipStreamSynchronize(c.stream));
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_update_positions(ljg_ctx *h, const double *x, const double *y, const double *z, const double *vx,
                                const double *vy, const double *vz, int32_t *moved) {
  if (!h || !x || !y || !z) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  if (const char *why = outside_open_region(c, c.n, x, y, z)) return fail(c, LJG_ERR_ARG, why);
  const bool vel = vx && vy && vz;
  const size_t n = (size_t)c.n;
  c.part_v.ensure((vel ? 6 : 3) * n);
  double *buf = c.part_v.p;
  const double *src[6] = {x, y, z, vx, vy, vz};
  for (int k = 0; k < (vel ? 6 : 3); ++k)
    LJG_HIP(hipMemcpyAsync(buf + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, c.stream, (int64_t)1, c.flag.p + 2);
  LJG_HIP(hipGetLastError());
  md_set_positions(c, buf, vel ? buf + 3 * n : nullptr);
  int32_t mv = 0;
  LJG_HIP(hipMemcpyAsync(&mv, c.flag.p + 2, sizeof mv, hipMemcpyDeviceToHost, c.stream));
  LJG_HIP(hipStreamSynchronize(c.stream));
  c.have_forces = false;
  c.box_glob_valid = false;
  if (moved) *moved = mv != 0;
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_rebuild(ljg_ctx *h, int32_t kernel) {
  if (!h || kernel < 0 || kernel > LJG_KERNEL_LAST) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  if (const char *why = kernel_unsupported(c, kernel)) return fail(c, LJG_ERR_UNSUPPORTED, why);
  rebuild_for(c, kernel);
  LJG_HIP(hipStreamSynchronize(c.stream));
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_bin(ljg_ctx *h) {
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  if (c.nboxes > 0) return fail(c, LJG_ERR_UNSUPPORTED, "linked cells on an ensemble of boxes");
  rebuild_for(c, LJG_KERNEL_LC_FP64);
  LJG_HIP(hipStreamSynchronize(c.stream));
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_build_pairs(ljg_ctx *h) {
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  rebuild_for(c, LJG_KERNEL_CLUSTER_MIXED);
  LJG_HIP(hipStreamSynchronize(c.stream));
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_force(ljg_ctx *h, int32_t kernel, int32_t calc_globals, ljg_globals *out) {
  if (!h || kernel < 0 || kernel > LJG_KERNEL_LAST) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  if (const char *why = kernel_unsupported(c, kernel)) return fail(c, LJG_ERR_UNSUPPORTED, why);
  if (calc_globals && is_n3_kernel(kernel) && owned_of(c).orig)
    return fail(c, LJG_ERR_UNSUPPORTED, "Newton-3 kernels count the globals of every particle: with halo particles "
                                        "use a full-list kernel");
  prepare_layout(c, kernel);
  run_force(c, kernel, calc_globals != 0);
  if (c.nboxes > 0 && calc_globals) ensemble_box_globals(c);
  c.box_glob_valid = c.nboxes > 0 && calc_globals;
  ljg_globals g{};
  reduce_partials(c, partial_items(c, kernel), &g, calc_globals != 0);
  g.candidate_evals = candidate_evals(c, kernel);
  if (!calc_globals) {
    g.potential_energy = 0.;
    g.virial = 0.;
  }
  c.have_forces = true;
  if (out) *out = g;
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_set_timing(ljg_ctx *h, int32_t enabled) {
  if (!h) return LJG_ERR_ARG;
  h->c.timing = enabled != 0;
  return LJG_OK;
}

// nsteps velocity-Verlet steps with one kernel (LogicHandler.h:987-1015 rebuild policy); the statistics
// are added to rs.  HIP-event timing of force launches and rebuilds when `timing`.
//
// Step s is: drift of s (fused with the kick of s - 1), a rebuild if one is due, force of s.  The steps
// between two scheduled rebuilds are queued as one chunk (at most kChunkSteps) and the host reads the
// displacement flag once at the chunk's end.  A drift that moves a particle skin/2 since the last rebuild
// raises the flag with its step's stamp; every later launch of the chunk exits at once (SkipIf), so the
// state stops right after that drift.  The host then rebuilds, clears the flag and resumes with that
// step's force, as LogicHandler.h:1000-1015 rebuilds before the force of the step that moved too far.
// The FP32 cluster kernels (j-row and cluster tiles) advance the particles in their epilogue (kick of step
// s, drift of step s + 1), so with them only the very first drift runs as a kernel of its own.
constexpr int32_t kTimingStride = 8;
constexpr int32_t kChunkSteps = 32;

// Clears the displacement flag and, when given, its host-mapped copy: in stream order, after every queued
// drift that may still raise it, so the host needs no synchronisation first.
__global__ void k_clear_flag(int32_t *flag, int32_t *host_copy) {
  flag[0] = 0;
  if (host_copy) host_copy[0] = 0;
}

// The pair counts of the step loop's kAccSlots lines, summed into *sum.
__global__ void k_sum_slots(const unsigned long long *__restrict__ slots, unsigned long long *__restrict__ sum) {
  if (threadIdx.x != 0) return;
  unsigned long long t = 0;
  for (int k = 0; k < kAccSlots; ++k) t += slots[k * kAccStride];
  sum[0] = t;
}

static hipEvent_t take_event(Ctx &c) {
  if (!c.event_pool.empty()) {
    hipEvent_t e = c.event_pool.back();
    c.event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  LJG_HIP(hipEventCreate(&e));
  return e;
}

static void run_steps(Ctx &c, int32_t kernel, int32_t nsteps, double dt, bool timing, ljg_run_stats &rs) {
  std::vector<hipEvent_t> rev;                           // pairs (start, stop) per rebuild
  std::vector<std::pair<int32_t, hipEvent_t>> fev;       // (step, event): start, stop per timed force launch
  c.hits_slots.ensure(kAccSlots * kAccStride + 1);  // (+1: their sum, read back at the end)
  unsigned long long *const sum_slots_out = c.hits_slots.p + kAccSlots * kAccStride;
  unsigned long long total_hits = 0;
  bool total_read = false;  // total_hits holds the run's sum, read with the last chunk's synchronisation
  hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(kAccSlots * kAccStride), 0, c.stream, (int64_t)(kAccSlots * kAccStride),
                     reinterpret_cast<int64_t *>(c.hits_slots.p));
  LJG_HIP(hipGetLastError());
  const bool saved_timing = c.timing;
  c.timing = false;  // per-launch events are recorded here instead
  const bool graphs = c.prm.graphs != 0;
  if (graphs) timing = false;  // no event records inside the captured chunks
  // clears both copies of the flag on the stream, after every queued launch (a drift that may still raise
  // it); the host reads its copy only after a later synchronisation
  auto clear_flag = [&]() {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, c.stream, c.flag.p, c.host_flag_dev);
    LJG_HIP(hipGetLastError());
    if (c.nboxes > 0) fill_i32(c, c.box_flag.p, c.nboxes, 0);
  };
  clear_flag();
  auto record_rebuild = [&]() {
    if (!timing) return;
    rev.push_back(take_event(c));
    LJG_HIP(hipEventRecord(rev.back(), c.stream));
  };
  record_rebuild();
  if (prepare_layout(c, kernel)) {
    record_rebuild();
    ++rs.rebuilds;
  } else if (timing) {
    c.event_pool.push_back(rev.back());
    rev.pop_back();
  }
  c.hits_acc = c.hits_slots.p;
  auto rebuild_now = [&]() {
    record_rebuild();
    c.carry_force = false;  // the force of this step is computed right after the rebuild
    rebuild_for(c, kernel);
    c.carry_force = true;
    record_rebuild();
    ++rs.rebuilds;
  };
  // the FP32 cluster kernels advance the particles in their epilogue
  const bool fused = kernel == LJG_KERNEL_CLUSTER_JROWS_MIXED || kernel == LJG_KERNEL_CLUSTER_MIXED;
  const bool jrows = kernel == LJG_KERNEL_CLUSTER_JROWS_MIXED;
  const bool dyn = c.prm.dynamic_rebuild != 0;
  // stamp of step s (s = 0 .. nsteps - 1): base + s + 1, never 0
  if (c.step_stamp > INT32_MAX - nsteps - 2) c.step_stamp = 0;  // both flags were just cleared
  const int32_t base = c.step_stamp;
  c.step_stamp += nsteps + 1;
  auto stamp_of = [&](int32_t step) { return base + step + 1; };
  // the force of step s; with `speculative`, it exits if the flag was raised at or before step s
  // (skip_after: with `speculative`, the stamp at or below which a raised flag makes it exit; default its
  // own step's)
  auto force_step = [&](int32_t step, bool speculative, int32_t skip_after = 0) {
    const bool time_force = timing && step % kTimingStride == 0;
    if (time_force) {
      fev.emplace_back(step, take_event(c));
      LJG_HIP(hipEventRecord(fev.back().second, c.stream));
    }
    c.skip = speculative && dyn ? SkipIf{c.flag.p, skip_after ? skip_after : stamp_of(step)} : SkipIf{nullptr, 0};
    if (fused) {
      const bool last = step + 1 == nsteps;
      const IntegArgs ia = integ_args(c, true, !last, dt, last ? 0 : stamp_of(step + 1));
      if (jrows) {
        cluster_jrows_force(c, false, &ia);
        if (!last) std::swap(c.posfx, c.posfx_next);  // the launch wrote the next step's positions there
      } else {
        cluster_force(c, true, false, &ia);
        if (!last) swap_cluster_positions(c);
      }
    } else {
      run_force(c, kernel, false);
    }
    c.skip = SkipIf{nullptr, 0};
    if (time_force) {
      fev.emplace_back(step, take_event(c));
      LJG_HIP(hipEventRecord(fev.back().second, c.stream));
    }
  };
  // the drift of step s (fused with the kick of s - 1); with a fused kernel only that of step 0
  auto drift_step = [&](int32_t step, bool speculative) {
    if (fused && step > 0) return;
    const SkipIf sk = speculative && dyn ? SkipIf{c.flag.p, stamp_of(step) - 1} : SkipIf{nullptr, 0};
    md_kick_drift(c, !fused && step > 0, true, dt, stamp_of(step), sk);
  };
  // An ensemble keeps a schedule per box, the one the box has alone: box b is rebuilt before the force of
  // the step at which its steps since its last rebuild reach the frequency, or whose drift moved one of
  // its particles skin/2 (it raises the box's flag with the shared one).  A raised flag stops every box
  // after that drift; only the boxes due then are rebuilt, in place, and all go on from that step.
  const bool per_box = c.nboxes > 0 && (int64_t)c.box_since.size() == c.nboxes && c.prm.shared_rebuilds == 0;
  std::vector<int32_t> mark;
  auto steps_to_due = [&]() {
    int64_t since = c.steps_since_rebuild;
    if (per_box) since = *std::max_element(c.box_since.begin(), c.box_since.end());
    return (int64_t)c.prm.rebuild_frequency - since;
  };
  int32_t s = 0;
  while (s < nsteps) {
    // steps on the current lists before the next scheduled rebuild, queued as one chunk
    int64_t k64 = steps_to_due();
    if (k64 < 0) k64 = 0;
    if (k64 > nsteps - s) k64 = nsteps - s;
    if (k64 > kChunkSteps) k64 = kChunkSteps;
    const int32_t k = (int32_t)k64;
    if (graphs && k > 0) {
      // the chunk's launches as one graph: captured from the stream, then the executable graph of the
      // previous chunk is updated with it when the launch sequence is the same, else instantiated anew
      LJG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
      try {
        for (int32_t j = 0; j < k; ++j) {
          drift_step(s + j, true);
          force_step(s + j, true);
        }
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hipGraph_t g = nullptr;
      LJG_HIP(hipStreamEndCapture(c.stream, &g));
      bool updated = false;
      if (c.graph_exec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(c.graph_exec, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) {
          (void)hipGetLastError();
          LJG_HIP(hipGraphExecDestroy(c.graph_exec));
          c.graph_exec = nullptr;
        }
      }
      if (!updated) {
        const hipError_t e = hipGraphInstantiate(&c.graph_exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          (void)hipGraphDestroy(g);
          c.graph_exec = nullptr;
          throw HipError{e, "hipGraphInstantiate", __LINE__};
        }
        ++c.graph_instantiations;
      }
      const hipError_t e = hipGraphLaunch(c.graph_exec, c.stream);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) throw HipError{e, "hipGraphLaunch", __LINE__};
      ++c.graph_launches;
    } else {
      for (int32_t j = 0; j < k; ++j) {
        drift_step(s + j, true);
        force_step(s + j, true);
      }
    }
    const bool due = s + k < nsteps && steps_to_due() - k <= 0;
    // A scheduled rebuild before the force of step s + k is queued at once, after that step's drift: it
    // works on the state where the queued steps stopped, which is also the state a raised flag asks to
    // rebuild at.  Otherwise the host reads the flag now.  (An ensemble reads the flags first: which
    // boxes to rebuild depends on where the steps stopped.)
    const bool rebuild_next = due && !per_box;
    // after a scheduled rebuild the force of step s + k is queued too, before the host waits: it exits if
    // the flag was raised at an earlier step, when the host rebuilds at that step instead
    const bool force_next = rebuild_next && fused;
    if (due) drift_step(s + k, true);
    // The boxes' flags come back with the same synchronisation: in the marker's words for up to
    // kMarkFlagBoxes boxes, else by a copy queued before it.
    const int flag_words = per_box && dyn && c.nboxes <= kMarkFlagBoxes ? (c.nboxes + 1) / 2 : 0;
    if (per_box && dyn) {
      if (c.pin_flags_cap < (size_t)c.nboxes + 1) {
        if (c.pin_flags) LJG_HIP(hipHostFree(c.pin_flags));
        c.pin_flags = nullptr;
        c.pin_flags_cap = 0;
        LJG_HIP(hipHostMalloc((void **)&c.pin_flags, ((size_t)c.nboxes + 1) * sizeof(int32_t), hipHostMallocDefault));
        c.pin_flags_cap = (size_t)c.nboxes + 1;
      }
      if (!flag_words)
        LJG_HIP(hipMemcpyAsync(c.pin_flags, c.box_flag.p, (size_t)c.nboxes * sizeof(int32_t), hipMemcpyDeviceToHost,
                               c.stream));
    }
    // The run ends with this chunk unless a flag was raised: its pair count comes back with the same
    // synchronisation.
    const bool last_chunk = !due && s + k == nsteps;
    bool sum_read = false;
    if (rebuild_next) {
      const unsigned long long marks = c.mark_seq;
      // a j-row build followed by its force leaves its list lengths to be read after that launch
      c.defer_lists = force_next && jrows;
      try {
        rebuild_now();
      } catch (...) {
        c.defer_lists = false;
        throw;
      }
      c.defer_lists = false;
      // the force after the build, checked: a launch on rows that overflowed exited, so build again
      auto force_after_build = [&]() {
        force_step(s + k, true, stamp_of(s + k) - 1);
        if (!vcl_finish_lists(c)) {
          rebuild_now();
          force_step(s + k, true, stamp_of(s + k) - 1);
        }
      };
      if (force_next && (!dyn || c.mark_seq != marks)) {
        // The build waited for the stream, so every drift queued before it has run and the host copy of
        // the flag is current.  Raised by no drift, or only by the one of step s + k (the rebuild just made
        // answers it): the force of step s + k runs, and the next chunk queues behind it without a wait.
        const int32_t early = dyn ? static_cast<volatile int32_t *>(c.host_flag)[0] : 0;
        if (early == 0 || early == stamp_of(s + k)) {
          if (early != 0) {
            static_cast<volatile int32_t *>(c.host_flag)[0] = 0;
            hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, c.stream, c.flag.p, (int32_t *)nullptr);
            LJG_HIP(hipGetLastError());
          }
          force_after_build();
          c.steps_since_rebuild = 1;
          for (auto &v : c.box_since) v = 1;
          total_read = false;
          s += k + 1;
          continue;
        }
      }
      if (force_next) force_after_build();
      stream_wait(c);
    } else if (dyn || due) {
      unsigned long long words[kMarkWords] = {};
      const int nsum = last_chunk ? 1 : 0;
      if (nsum) {
        hipLaunchKernelGGL(k_sum_slots, dim3(1), dim3(64), 0, c.stream, c.hits_slots.p, sum_slots_out);
        LJG_HIP(hipGetLastError());
      }
      stream_wait(c, reinterpret_cast<const unsigned long long *>(c.box_flag.p), flag_words, words,
                  nsum ? sum_slots_out : nullptr, nsum);
      if (flag_words) std::memcpy(c.pin_flags, words, (size_t)c.nboxes * sizeof(int32_t));
      if (nsum) {
        total_hits = words[flag_words];
        sum_read = true;
      }
    }
    const int32_t raised = dyn ? static_cast<volatile int32_t *>(c.host_flag)[0] : 0;
    // (the sum read with this chunk's synchronisation holds unless more steps follow)
    total_read = sum_read && raised == 0;
    if (force_next && (raised == 0 || raised >= stamp_of(s + k))) {
      // The queued force of step s + k ran.  A flag raised by the drift of step s + k asked for the
      // rebuild that preceded it: cleared (nothing is queued, so no wait).  One raised by that force's own
      // drift of step s + k + 1 stays for the next chunk, whose launches then exit until it is handled.
      if (raised == stamp_of(s + k)) {
        static_cast<volatile int32_t *>(c.host_flag)[0] = 0;
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, c.stream, c.flag.p, (int32_t *)nullptr);
        LJG_HIP(hipGetLastError());
      }
      c.steps_since_rebuild = 1;
      for (auto &v : c.box_since) v = 1;
      s += k + 1;
      continue;
    }
    if (force_next && s + k + 1 != nsteps) {
      // the queued force exited without writing the next positions: undo its swap of the buffers
      if (jrows) std::swap(c.posfx, c.posfx_next);
      else swap_cluster_positions(c);
    }
    if (raised == 0 && !due) {
      s += k;
      c.steps_since_rebuild += k;
      for (auto &v : c.box_since) v += k;
      continue;
    }
    // step t is next: steps s .. t - 1 are complete and step t's drift is done; a raised flag names t
    // (the drift that moved a particle skin/2), and nothing queued after that drift ran
    const int32_t t = raised != 0 ? raised - base - 1 : s + k;
    if (t < s || t > s + k || (!fused && !due && t == s + k))
      throw HipError{hipErrorUnknown, "displacement flag outside the queued steps", __LINE__};
    // (the rebuild recomputes the fixed-point positions from the FP64 state, so the position buffers the
    // skipped launches swapped need no repair)
    for (auto it = fev.begin(); it != fev.end();) {  // timed launches of skipped steps
      if (it->first >= t) {
        c.event_pool.push_back(it->second);
        it = fev.erase(it);
      } else {
        ++it;
      }
    }
    if (per_box) {
      // The force launches of steps t .. s + k - 1 exited at once, yet each swapped the position buffers
      // (all but the run's last step): undo an odd count, since the boxes not rebuilt read theirs on.
      if (fused && !jrows) {
        int32_t swaps = 0;
        for (int32_t j = t; j < s + k; ++j) swaps += j + 1 != nsteps;
        if (swaps & 1) swap_cluster_positions(c);
      }
      // the boxes whose drift of step t raised the flag, and those due at step t
      mark.assign((size_t)c.nboxes, 0);
      bool all = true;
      for (int32_t b = 0; b < c.nboxes; ++b) {
        c.box_since[b] += t - s;
        mark[b] = (raised != 0 && c.pin_flags[b] != 0) || c.box_since[b] >= c.prm.rebuild_frequency;
        all = all && mark[b];
      }
      if (fused && !jrows && !all && c.regions_valid && layout_matches(c, kernel) && prepare_segments(c, mark)) {
        // The marked boxes are rebuilt on the second stream while the force of step t runs for all other
        // boxes; then the force of the marked boxes.  The two sets share no cluster, slot or list row.
        // Everything queued has run (the chunk's synchronisation), so the flags clear without waiting.
        static_cast<volatile int32_t *>(c.host_flag)[0] = 0;
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, c.stream, c.flag.p, (int32_t *)nullptr);
        LJG_HIP(hipGetLastError());
        fill_i32(c, c.box_flag.p, c.nboxes, 0);
        const bool last = t + 1 == nsteps;
        const IntegArgs ia = integ_args(c, true, !last, dt, last ? 0 : stamp_of(t + 1));
        const int64_t nb_o = cluster_force_blocks(c, c.seg_nothers), nb_m = cluster_force_blocks(c, c.seg_nclusters);
        c.part_e.ensure((size_t)(nb_o + nb_m));
        c.part_v.ensure((size_t)(nb_o + nb_m));
        c.part_hits.ensure((size_t)(nb_o + nb_m));
        if (!c.side) LJG_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
        if (!c.ev_fork) LJG_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
        if (!c.ev_join) LJG_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
        LJG_HIP(hipEventRecord(c.ev_fork, c.stream));
        LJG_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
        c.skip = SkipIf{nullptr, 0};
        const ClusterSubset others{c.seg_others, c.seg_nothers, 0};
        cluster_force(c, true, false, &ia, &others);
        const hipStream_t main_stream = c.stream;
        c.stream = c.side;
        try {
          record_rebuild();
          c.carry_force = false;
          rebuild_boxes(c, kernel, mark, true);
          c.carry_force = true;
          record_rebuild();
          const ClusterSubset marked{c.seg_clusters, c.seg_nclusters, nb_o};
          cluster_force(c, true, false, &ia, &marked);
          LJG_HIP(hipEventRecord(c.ev_join, c.side));
        } catch (...) {
          c.stream = main_stream;
          c.carry_force = true;
          (void)hipStreamSynchronize(c.side);
          throw;
        }
        c.stream = main_stream;
        LJG_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
        if (!last) swap_cluster_positions(c);
        ++rs.rebuilds;
        c.steps_since_rebuild = 1;
        for (auto &v : c.box_since) v += 1;
        s = t + 1;
        continue;
      }
      record_rebuild();
      c.carry_force = false;
      rebuild_boxes(c, kernel, mark);
      c.carry_force = true;
      record_rebuild();
      ++rs.rebuilds;
    } else if (!rebuild_next) {
      rebuild_now();
    }
    clear_flag();
    force_step(t, false);
    c.steps_since_rebuild = 1;
    for (auto &v : c.box_since) v += 1;
    s = t + 1;
  }
  c.hits_acc = nullptr;
  if (nsteps > 0 && !fused) {
    md_kick_drift(c, true, false, dt, 0);  // final velocity update
    total_read = false;
  }
  c.timing = saved_timing;
  if (!total_read) {
    hipLaunchKernelGGL(k_sum_slots, dim3(1), dim3(64), 0, c.stream, c.hits_slots.p, sum_slots_out);
    LJG_HIP(hipGetLastError());
    stream_wait(c, sum_slots_out, 1, &total_hits);
  }
  const int64_t hits2 = (int64_t)total_hits;
  rs.steps += nsteps;
  rs.pair_interactions += hits2 / 2;
  for (size_t k = 0; k + 1 < fev.size(); k += 2) {
    float ms = 0.f;
    LJG_HIP(hipEventElapsedTime(&ms, fev[k].second, fev[k + 1].second));
    rs.force_ms += ms;
    rs.force_launches++;
  }
  for (size_t k = 0; k + 1 < rev.size(); k += 2) {
    float ms = 0.f;
    LJG_HIP(hipEventElapsedTime(&ms, rev[k], rev[k + 1]));
    rs.rebuild_ms += ms;
  }
  for (auto &e : fev) c.event_pool.push_back(e.second);
  for (auto e : rev) c.event_pool.push_back(e);
  c.have_forces = true;
}


// ---- layout tuning (AutoTuner.cpp:85-175,250-369 in spirit) ----------------------------------------
// Candidates: FP32-pair/FP64-accumulate layouts (kernel, cluster size, i-clusters per wavefront of the
// force kernel, of the j-row expansion pass).  A tuning phase runs each candidate the box supports for
// one unscored step (it absorbs the candidate's one-time costs: the first launches of its kernels and the
// growth of its buffers, which AutoPas's tuner spreads over many more iterations) and then tuning_samples
// scored steps of the real simulation, the first of them after a rebuild, and scores it as AutoTuner's
// estimateRuntimeFromSamples does with the mean selector: rebuild time / rebuild_frequency + the rest of
// the step.  The lowest score is used until tuning_interval steps later (0: never again).
struct TuneCandidate {
  int32_t kernel, M, G, B;
};
static const TuneCandidate kTuneCandidates[] = {
    {LJG_KERNEL_CLUSTER_JROWS_MIXED, 8, 8, 4}, {LJG_KERNEL_CLUSTER_JROWS_MIXED, 8, 4, 4},
    {LJG_KERNEL_CLUSTER_JROWS_MIXED, 4, 8, 4}, {LJG_KERNEL_CLUSTER_JROWS_MIXED, 4, 16, 4},
    {LJG_KERNEL_CLUSTER_MIXED, 4, 8, 0},       {LJG_KERNEL_CLUSTER_MIXED, 8, 2, 0},
    {LJG_KERNEL_CLUSTER_MIXED, 8, 8, 0},       {LJG_KERNEL_CLUSTER_MIXED_N3, 8, 2, 0},
    {LJG_KERNEL_VERLET_MIXED, 4, 0, 0},
};
constexpr int kNumTuneCandidates = (int)(sizeof(kTuneCandidates) / sizeof(kTuneCandidates[0]));
static_assert(kNumTuneCandidates <= LJG_MAX_TUNE_CANDIDATES, "ljg_tuning holds every candidate");

static void apply_candidate(Ctx &c, const TuneCandidate &k) {
  c.M = k.M;
  c.prm.tile_groups = k.G;
  c.prm.build_groups = k.B;
}

static int next_candidate(const Ctx &c, int from) {
  for (int i = from; i < kNumTuneCandidates; ++i)
    if (!kernel_unsupported(c, kTuneCandidates[i].kernel)) return i;
  return -1;
}

static void run_auto(Ctx &c, int32_t nsteps, double dt, ljg_run_stats &rs) {
  TuneState &t = c.tune;
  const int32_t samples = c.prm.tuning_samples > 0 ? c.prm.tuning_samples
                                                    : (c.prm.rebuild_frequency > 3 ? c.prm.rebuild_frequency : 3);
  int32_t left = nsteps;
  while (left > 0) {
    if (!t.active && (t.best < 0 || (c.prm.tuning_interval > 0 && t.steps_since >= c.prm.tuning_interval))) {
      t.active = true;
      t.ci = next_candidate(c, 0);
      t.done = 0;
      t.warmed = false;
      t.wall_ms = t.rebuild_ms = 0.;
      t.rebuilds = 0;
      for (double &e : t.effective_ms) e = -1.;
      ++t.phases;
    }
    if (t.active) {
      const TuneCandidate &k = kTuneCandidates[t.ci];
      apply_candidate(c, k);
      if (!t.warmed) {
        ljg_run_stats sub{};
        run_steps(c, k.kernel, 1, dt, false, sub);
        t.warmed = true;
        --left;
        rs.steps += sub.steps;
        rs.rebuilds += sub.rebuilds;
        rs.pair_interactions += sub.pair_interactions;
        continue;
      }
      // every candidate's samples start with a rebuild, so each score carries its rebuild cost
      if (t.done == 0) c.steps_since_rebuild = c.prm.rebuild_frequency;
      const int32_t run = left < samples - t.done ? left : samples - t.done;
      ljg_run_stats sub{};
      const auto t0 = std::chrono::steady_clock::now();
      run_steps(c, k.kernel, run, dt, true, sub);  // ends with a synchronisation
      t.wall_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      t.rebuild_ms += sub.rebuild_ms;
      t.rebuilds += (int32_t)sub.rebuilds;
      t.done += run;
      left -= run;
      rs.steps += sub.steps;
      rs.rebuilds += sub.rebuilds;
      rs.pair_interactions += sub.pair_interactions;
      if (c.timing) {
        rs.force_ms += sub.force_ms;
        rs.force_launches += sub.force_launches;
        rs.rebuild_ms += sub.rebuild_ms;
      }
      if (t.done == samples) {
        const double per_rebuild = t.rebuilds > 0 ? t.rebuild_ms / t.rebuilds : 0.;
        t.effective_ms[t.ci] = (t.wall_ms - t.rebuild_ms) / samples + per_rebuild / c.prm.rebuild_frequency;
        const int nxt = next_candidate(c, t.ci + 1);
        if (nxt >= 0) {
          t.ci = nxt;
          t.done = 0;
          t.warmed = false;
          t.wall_ms = t.rebuild_ms = 0.;
          t.rebuilds = 0;
        } else {
          int best = -1;
          for (int i = 0; i < kNumTuneCandidates; ++i)
            if (t.effective_ms[i] >= 0. && (best < 0 || t.effective_ms[i] < t.effective_ms[best])) best = i;
          t.best = best;
          t.active = false;
          t.steps_since = 0;
        }
      }
    } else {
      const TuneCandidate &k = kTuneCandidates[t.best];
      apply_candidate(c, k);
      int32_t run = left;
      if (c.prm.tuning_interval > 0 && c.prm.tuning_interval - t.steps_since < run)
        run = (int32_t)(c.prm.tuning_interval - t.steps_since);
      run_steps(c, k.kernel, run, dt, c.timing, rs);
      t.steps_since += run;
      left -= run;
    }
  }
}

ljg_status ljg_step(ljg_ctx *h, int32_t nsteps, double dt, int32_t kernel, ljg_run_stats *out) {
  if (!h || nsteps < 0 || kernel < 0 || (kernel > LJG_KERNEL_LAST && kernel != LJG_KERNEL_AUTO) || !(dt > 0.))
    return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  if (c.open)
    return fail(c, LJG_ERR_UNSUPPORTED, "an open box takes one force evaluation at a time: the caller moves the "
                                        "particles and exchanges halos between them");
  if (kernel != LJG_KERNEL_AUTO)
    if (const char *why = kernel_unsupported(c, kernel)) return fail(c, LJG_ERR_UNSUPPORTED, why);
  const auto t0 = std::chrono::steady_clock::now();
  ljg_run_stats rs{};
  c.box_glob_valid = false;
  if (kernel == LJG_KERNEL_AUTO) {
    if (next_candidate(c, 0) < 0) return fail(c, LJG_ERR_UNSUPPORTED, "no tuning candidate fits this box");
    run_auto(c, nsteps, dt, rs);
  } else {
    run_steps(c, kernel, nsteps, dt, c.timing, rs);
  }
  rs.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (out) *out = rs;
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_get_tuning(ljg_ctx *h, ljg_tuning *out) {
  if (!h || !out) return LJG_ERR_ARG;
  Ctx &c = h->c;
  std::memset(out, 0, sizeof *out);
  out->candidates = kNumTuneCandidates;
  out->best = c.tune.best;
  out->phases = c.tune.phases;
  out->tuning = c.tune.active ? 1 : 0;
  for (int i = 0; i < kNumTuneCandidates; ++i) {
    out->kernel[i] = kTuneCandidates[i].kernel;
    out->cluster_size[i] = kTuneCandidates[i].M;
    out->tile_groups[i] = kTuneCandidates[i].G;
    out->build_groups[i] = kTuneCandidates[i].B;
    out->effective_ms[i] = c.tune.effective_ms[i];
  }
  return LJG_OK;
}

ljg_status ljg_download(ljg_ctx *h, double *x, double *y, double *z, double *vx, double *vy, double *vz, double *fx,
                        double *fy, double *fz) {
  LJG_API_BEGIN(h)
  if (c.n == 0) return fail(c, LJG_ERR_STATE, "no particles uploaded");
  gather_range(c, 0, c.n, x, y, z, vx, vy, vz, fx, fy, fz);
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_get_cells(ljg_ctx *h, int64_t *cell_of_particle, int64_t nh[3]) {
  LJG_API_BEGIN(h)
  if (c.layout != Layout::Cells) return fail(c, LJG_ERR_STATE, "particles are not binned into cells");
  if (cell_of_particle)
    LJG_HIP(hipMemcpy(cell_of_particle, c.cell_of_particle.p, (size_t)c.n * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (nh)
    for (int d = 0; d < 3; ++d) nh[d] = c.cg.nh[d];
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_cluster_sizes(ljg_ctx *h, int64_t *num_slots, int64_t *num_clusters, int64_t *num_entries,
                             int64_t towers_per_dim[2], double side[2], int64_t *towers_per_il) {
  LJG_API_BEGIN(h)
  if (c.layout != Layout::Clusters) return fail(c, LJG_ERR_STATE, "no cluster structure built");
  if (num_slots) *num_slots = c.nslots;
  if (num_clusters) *num_clusters = c.nclusters;
  if (num_entries) *num_entries = c.nentries;
  if (towers_per_dim) { towers_per_dim[0] = c.tg.nt[0]; towers_per_dim[1] = c.tg.nt[1]; }
  if (side) { side[0] = c.tg.side[0]; side[1] = c.tg.side[1]; }
  if (towers_per_il) *towers_per_il = c.tg.R;
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_get_clusters(ljg_ctx *h, int64_t *slot_particle, int32_t *cluster_tower, int64_t *list_offsets,
                            int32_t *list_j, int32_t *list_shift, double *bbox) {
  LJG_API_BEGIN(h)
  if (c.layout != Layout::Clusters) return fail(c, LJG_ERR_STATE, "no cluster structure built");
  LJG_HIP(hipStreamSynchronize(c.stream));
  if (slot_particle) {
    std::vector<int32_t> o((size_t)c.nslots);
    LJG_HIP(hipMemcpy(o.data(), c.st[c.cur].orig.p, o.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < o.size(); ++k) slot_particle[k] = o[k];
  }
  if (cluster_tower)
    LJG_HIP(hipMemcpy(cluster_tower, c.cluster_tower.p, (size_t)c.nclusters * sizeof(int32_t), hipMemcpyDeviceToHost));
  // lists live in fixed-capacity rows on the device; hand them out compacted
  std::vector<int32_t> cnt((size_t)c.nclusters);
  if (!cnt.empty())
    LJG_HIP(hipMemcpy(cnt.data(), c.jrows_valid ? c.jcount.p : c.list_count.p, cnt.size() * sizeof(int32_t),
                      hipMemcpyDeviceToHost));
  if (list_offsets) {
    list_offsets[0] = 0;
    for (size_t k = 0; k < cnt.size(); ++k) list_offsets[k + 1] = list_offsets[k] + cnt[k];
  }
  if ((list_j || list_shift) && c.jrows_valid) {
    // j-particle rows: upload index of each j, no shift (minimum image)
    std::vector<int32_t> l((size_t)c.nclusters * c.list_row), o((size_t)c.nslots);
    if (!l.empty()) LJG_HIP(hipMemcpy(l.data(), c.jlist.p, l.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    LJG_HIP(hipMemcpy(o.data(), c.st[c.cur].orig.p, o.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    size_t w = 0;
    for (size_t ci = 0; ci < cnt.size(); ++ci) {
      for (int32_t k = 0; k < cnt[ci]; ++k, ++w) {
        if (list_j) list_j[w] = o[l[ci * c.list_row + k]];
        if (list_shift) list_shift[w] = -1;
      }
    }
  } else if (list_j || list_shift) {
    std::vector<int4> l((size_t)c.nclusters * c.list_row);
    if (!l.empty()) LJG_HIP(hipMemcpy(l.data(), c.list.p, l.size() * sizeof(int4), hipMemcpyDeviceToHost));
    size_t w = 0;
    for (size_t ci = 0; ci < cnt.size(); ++ci) {
      for (int32_t k = 0; k < cnt[ci]; ++k, ++w) {
        const int4 &en = l[ci * c.list_row + k];
        if (list_j) list_j[w] = en.x & kClusterMask;
        if (list_shift) list_shift[w] = (int32_t)((uint32_t)en.x >> kShiftBits);
      }
    }
  }
  if (bbox) LJG_HIP(hipMemcpy(bbox, c.bbox.p, (size_t)c.nclusters * 6 * sizeof(double), hipMemcpyDeviceToHost));
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_get_member_rows(ljg_ctx *h, int64_t *offsets, int32_t *rows) {
  LJG_API_BEGIN(h)
  if (c.layout != Layout::Clusters || !c.xrows_valid) return fail(c, LJG_ERR_STATE, "no member rows built");
  if (!offsets) return fail(c, LJG_ERR_ARG, "offsets is required");
  LJG_HIP(hipStreamSynchronize(c.stream));
  const size_t ncl = (size_t)c.nclusters;
  std::vector<int32_t> cnt(ncl);
  if (ncl > 0) LJG_HIP(hipMemcpy(cnt.data(), c.xcount.p, ncl * sizeof(int32_t), hipMemcpyDeviceToHost));
  offsets[0] = 0;
  for (size_t k = 0; k < ncl; ++k) offsets[k + 1] = offsets[k] + cnt[k];
  if (rows && ncl > 0) {
    std::vector<int2> r(ncl * (size_t)c.xstride);
    LJG_HIP(hipMemcpy(r.data(), c.xrow.p, r.size() * sizeof(int2), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ncl; ++k)
      for (int32_t i = 0; i < cnt[k]; ++i) {
        const int2 v = r[k * (size_t)c.xstride + i];
        rows[offsets[k] + i] = (v.y << 3) | (v.x % c.M);
      }
  }
  return LJG_OK;
  LJG_API_END
}

ljg_status ljg_get_verlet_lists(ljg_ctx *h, int64_t *offsets, int32_t *neighbors) {
  if (!h || !offsets) return LJG_ERR_ARG;
  LJG_API_BEGIN(h)
  if (c.layout != Layout::Verlet) return fail(c, LJG_ERR_STATE, "no Verlet lists built");
  const int64_t n = c.n;
  std::vector<int32_t> cnt((size_t)n), orig((size_t)n);
  LJG_HIP(hipMemcpy(cnt.data(), c.vcount.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  LJG_HIP(hipMemcpy(orig.data(), c.st[c.cur].orig.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<int64_t> slot_of((size_t)n);
  for (int64_t s = 0; s < n; ++s) slot_of[orig[s]] = s;
  offsets[0] = 0;
  for (int64_t p = 0; p < n; ++p) offsets[p + 1] = offsets[p] + cnt[slot_of[p]];
  if (neighbors) {
    const int64_t nw = (n + kWave - 1) / kWave;
    std::vector<int32_t> l((size_t)(nw * c.vrow * kWave));
    LJG_HIP(hipMemcpy(l.data(), c.vlist.p, l.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < n; ++p) {
      const int64_t s = slot_of[p], w = s / kWave, lane = s % kWave;
      for (int32_t k = 0; k < cnt[s]; ++k) neighbors[offsets[p] + k] = orig[l[(w * c.vrow + k) * kWave + lane]];
    }
  }
  return LJG_OK;
  LJG_API_END
}

}  // extern "C"
