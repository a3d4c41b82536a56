// ljg_noalloc.hip -- ljg_force_lc_noalloc: the plain FP64 LinkedCells force evaluation of ljg_upload +
// ljg_force(LJG_KERNEL_LC_FP64) on caller-owned device buffers, with every intermediate in one static
// device array, so that the call allocates nothing (for callers under HIP graph capture or in loops that
// must not allocate).  The same binning, id order, force kernel and reduction run as in a context, so the
// forces, energy, virial and pair count are bit-identical to the context's.
//
// Compiled with -ffp-contract=off (the wrap and the cell index must round as in ljg_bin.hip / ljg_md.hip).
#include <mutex>

#include "ljg_cells.h"
#include "ljg_internal.h"

namespace ljg {

// The scratch array: one copy per GPU, held by the library until the process ends.  Host code never uses
// the name itself as an address: each device's copy is looked up with hipGetSymbolAddress once, with that
// device current, and only that address goes into kernel arguments.
static __device__ __align__(256) unsigned char scratch[LJG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_setup;
unsigned char *g_scratch[kMaxDevices] = {};

// Byte offsets of the buffers inside the scratch array (each 256-aligned).
struct ScratchLayout {
  size_t cell_of, rank, items, sorted, cells, type, xw, xs, count, start, scan, pe, pv, ph, red, red_h, tab, total;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

ScratchLayout layout_for(int64_t n, int64_t ncells, int32_t ntypes, int64_t scan_len, int64_t nblocks) {
  ScratchLayout l{};
  size_t at = 0;
  auto take = [&at](size_t bytes) {
    const size_t o = at;
    at += align256(bytes);
    return o;
  };
  const size_t un = (size_t)n;
  l.cell_of = take(4 * un);
  l.rank = take(4 * un);
  l.items = take(4 * un);
  l.sorted = take(4 * un);
  l.cells = take(4 * un);
  l.type = take(4 * un);
  l.xw = take(24 * un);  // wrapped positions in caller order, later the forces in cell order
  l.xs = take(24 * un);  // positions in cell order
  l.count = take(4 * (size_t)(ncells + 1));
  l.start = take(8 * (size_t)(ncells + 1));
  l.scan = take(8 * (size_t)scan_len);
  l.pe = take(8 * (size_t)nblocks);
  l.pv = take(8 * (size_t)nblocks);
  l.ph = take(8 * (size_t)nblocks);
  l.red = take(16);
  l.red_h = take(8);
  l.tab = take(24 * (size_t)ntypes * ntypes);
  l.total = at;
  return l;
}

__global__ void k_na_fill(int64_t n, int32_t *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Up to 256 doubles of the mixing table, passed by value (no host-to-device copy).
struct TabChunk {
  int32_t count;
  double v[256];
};
__global__ void k_na_tab(double *__restrict__ dst, TabChunk c) {
  const int i = threadIdx.x;
  if (i < c.count) dst[i] = c.v[i];
}

// Positions wrapped into the box as ljg_upload wraps them (RegularGridDecomposition.cpp:563-587).
__global__ void k_na_wrap(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                          const double *__restrict__ z, WrapArgs w, double *__restrict__ xw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xw[i] = wrap_coord(x[i], 0, w);
  xw[n + i] = wrap_coord(y[i], 1, w);
  xw[2 * n + i] = wrap_coord(z[i], 2, w);
}

// Slot s of the cell order takes particle sorted[s]: its wrapped position, type and cell.  Types are
// clamped to [0, ntypes) so that a bad type cannot index past the mixing table.
__global__ void k_na_gather(int64_t n, const int32_t *__restrict__ sorted, const double *__restrict__ xw,
                            const int32_t *__restrict__ type, int32_t ntypes, const int32_t *__restrict__ cell_of,
                            double *__restrict__ xs, int32_t *__restrict__ ts, int32_t *__restrict__ cells) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int32_t k = sorted[s];
  xs[s] = xw[k];
  xs[n + s] = xw[n + k];
  xs[2 * n + s] = xw[2 * n + k];
  const int32_t t = type ? type[k] : 0;
  ts[s] = t < 0 ? 0 : t >= ntypes ? ntypes - 1 : t;
  cells[s] = cell_of[k];
}

// Forces of the cell order back to the caller's order.
__global__ void k_na_scatter(int64_t n, const int32_t *__restrict__ sorted, const double *__restrict__ fs,
                             double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int32_t k = sorted[s];
  fx[k] = fs[s];
  fy[k] = fs[n + s];
  fz[k] = fs[2 * n + s];
}

// LJFunctor::endTraversal's scaling of the full-shell sums (LJFunctor.h:661-685), as ljg_force applies it.
__global__ void k_na_globals(const double *__restrict__ red, const int64_t *__restrict__ red_h,
                             ljg_globals *__restrict__ out) {
  if (threadIdx.x != 0) return;
  out->potential_energy = red[0] * 0.5 / 6.;
  out->virial = red[1] * 0.5;
  out->pair_count = red_h[0] / 2;
  out->candidate_evals = -1;
}

}  // namespace
}  // namespace ljg

using namespace ljg;

extern "C" {

ljg_status ljg_force_lc_noalloc(const ljg_params *params, int32_t ntypes, const double *epsilon, const double *sigma,
                                int64_t n, const double *x, const double *y, const double *z, const int32_t *type,
                                const int64_t *id, double *fx, double *fy, double *fz, ljg_globals *globals,
                                void *stream) {
  if (!params || n <= 0 || !x || !y || !z || !fx || !fy || !fz) return LJG_ERR_ARG;
  if (ntypes < 1 || ntypes > kMaxTypes || (ntypes > 1 && (!epsilon || !sigma)))
    return LJG_ERR_ARG;
  const double ilen = params->cutoff + params->skin;
  if (!(params->cutoff > 0.) || !(params->skin >= 0.)) return LJG_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (!(params->box_max[d] - params->box_min[d] >= ilen)) return LJG_ERR_ARG;
  if (n > LJG_NOALLOC_MAX_PARTICLES) return LJG_ERR_CAPACITY;
  int64_t ncells = 0;
  const CellGeom g = make_cell_geom(params->box_min, params->box_max, ilen, &ncells);
  if (ncells > LJG_NOALLOC_MAX_CELLS) return LJG_ERR_CAPACITY;
  const int64_t nblocks = blocks_for(n, 256);
  const int64_t slen = scan_tmp_len(ncells);
  const ScratchLayout l = layout_for(n, ncells, ntypes, slen, nblocks);
  if (l.total > (size_t)LJG_NOALLOC_SCRATCH_BYTES) return LJG_ERR_CAPACITY;

  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return LJG_ERR_NODEVICE;
  if (dev < 0 || dev >= kMaxDevices) return LJG_ERR_UNSUPPORTED;
  unsigned char *base = nullptr;
  {
    // set-up once per device, with that device current: the address of this device's copy
    std::lock_guard<std::mutex> lock(g_setup);
    if (!g_scratch[dev]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return LJG_ERR_HIP;
      g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    base = g_scratch[dev];
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  int32_t *cell_of = reinterpret_cast<int32_t *>(base + l.cell_of);
  int32_t *rank = reinterpret_cast<int32_t *>(base + l.rank);
  int32_t *items = reinterpret_cast<int32_t *>(base + l.items);
  int32_t *sorted = reinterpret_cast<int32_t *>(base + l.sorted);
  int32_t *cells = reinterpret_cast<int32_t *>(base + l.cells);
  int32_t *ts = reinterpret_cast<int32_t *>(base + l.type);
  double *xw = reinterpret_cast<double *>(base + l.xw);
  double *xs = reinterpret_cast<double *>(base + l.xs);
  int32_t *count = reinterpret_cast<int32_t *>(base + l.count);
  int64_t *start = reinterpret_cast<int64_t *>(base + l.start);
  int64_t *scan = reinterpret_cast<int64_t *>(base + l.scan);
  double *pe = reinterpret_cast<double *>(base + l.pe);
  double *pv = reinterpret_cast<double *>(base + l.pv);
  int64_t *ph = reinterpret_cast<int64_t *>(base + l.ph);
  double *red = reinterpret_cast<double *>(base + l.red);
  int64_t *red_h = reinterpret_cast<int64_t *>(base + l.red_h);
  double *tab = reinterpret_cast<double *>(base + l.tab);

  // the mixing table (FP64, as ljg_set_types forms it)
  const double rc2 = params->cutoff * params->cutoff;
  double e24[kMaxTypes * kMaxTypes], s2[kMaxTypes * kMaxTypes], sh6[kMaxTypes * kMaxTypes];
  const double one = 1.0;  // one type without parameters: epsilon = sigma = 1, as a new context has
  mix_table(ntypes, epsilon ? epsilon : &one, sigma ? sigma : &one, rc2, params->apply_shift != 0, e24, s2, sh6);
  try {
    if (ntypes > 1) {
      const int32_t nt = 3 * ntypes * ntypes;
      for (int32_t off = 0; off < nt; off += 256) {
        TabChunk ch{};
        ch.count = nt - off < 256 ? nt - off : 256;
        for (int32_t k = 0; k < ch.count; ++k) {
          const int32_t q = off + k;
          ch.v[k] = q % 3 == 0 ? e24[q / 3] : q % 3 == 1 ? s2[q / 3] : sh6[q / 3];
        }
        hipLaunchKernelGGL(k_na_tab, dim3(1), dim3(256), 0, st, tab + off, ch);
      }
    }
    // every counter that must start at zero is cleared here by a kernel (the array keeps the last call's
    // contents)
    hipLaunchKernelGGL(k_na_fill, dim3(blocks_for(ncells + 1, 256)), dim3(256), 0, st, ncells + 1, count);
    hipLaunchKernelGGL(k_na_wrap, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, x, y, z, wrap_args(*params), xw);
    LJG_HIP(hipGetLastError());
    cell_bin_launch(st, n, xw, xw + n, xw + 2 * n, g, cell_of, rank, count, nullptr, nullptr, nullptr, nullptr, 0);
    exclusive_scan_i32_to_i64_on(st, count, start, ncells, scan, slen);
    cell_bin_launch(st, n, nullptr, nullptr, nullptr, g, cell_of, rank, count, start, items, id, sorted, 1);
    hipLaunchKernelGGL(k_na_gather, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, sorted, xw, type, ntypes, cell_of, xs,
                       ts, cells);
    LJG_HIP(hipGetLastError());
    // the forces in cell order go where the wrapped caller-order positions were
    double *fs = xw;
    lc_force_launch(st, n, xs, xs + n, xs + 2 * n, ts, cells, start, g, rc2, ntypes, tab, e24[0], s2[0], sh6[0], fs,
                    fs + n, fs + 2 * n, pe, pv, ph, globals != nullptr, nullptr, SkipIf{nullptr, 0});
    hipLaunchKernelGGL(k_na_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, sorted, fs, fx, fy, fz);
    LJG_HIP(hipGetLastError());
    if (globals) {
      reduce_launch(st, nblocks, pe, pv, ph, red, red_h, true);
      hipLaunchKernelGGL(k_na_globals, dim3(1), dim3(64), 0, st, red, red_h, globals);
      LJG_HIP(hipGetLastError());
    }
  } catch (const HipError &e) {
    return LJG_ERR_HIP;
  }
  return LJG_OK;
}

}  // extern "C"
