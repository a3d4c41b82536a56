// ljg_bin.hip -- LinkedCells binning on the device: position -> CellBlock3D cell (bit-exact restatement
// of CellBlock3D::get3DIndexOfPosition), counting sort (histogram, exclusive scan, scatter), a per-cell
// stable order by particle id, and one gather pass that writes the whole particle state in cell order
// (coalesced SoA).  Compiled with -ffp-contract=off: the index arithmetic must round exactly as the
// reference's FP64 expression.
#include <cmath>
#include <utility>

#include "ljg_cells.h"
#include "ljg_internal.h"

namespace ljg {

__global__ void k_cell_hist(int64_t nslots, const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const int32_t *__restrict__ orig, CellGeom g,
                            int32_t *__restrict__ cell_of, int32_t *__restrict__ rank,
                            int32_t *__restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < nslots && (orig == nullptr || orig[i] >= 0);
  const int32_t c = active ? cell_index_1d(g, x[i], y[i], z[i]) : -1;
  const int32_t r = wave_aggregated_rank(c, active, count);
  if (i < nslots) {
    cell_of[i] = c;
    rank[i] = r;
  }
}

__global__ void k_cell_scatter(int64_t nslots, const int32_t *__restrict__ cell_of, const int32_t *__restrict__ rank,
                               const int64_t *__restrict__ start, int32_t *__restrict__ items) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  const int32_t c = cell_of[i];
  if (c < 0) return;
  items[start[c] + rank[i]] = (int32_t)i;
}

// The atomic histogram hands out ranks in arbitrary order; every cell's members are put in particle-id
// order so the layout (and with it every floating-point summation order downstream) is reproducible.
// A thread per member counts the members of its cell with a smaller id (ids are distinct) and writes
// itself to that place: independent loads of the cell's few entries, instead of a serial sort per cell.
__global__ void k_cell_order_by_id(int64_t n, const int32_t *__restrict__ items, const int32_t *__restrict__ cell_of,
                                   const int64_t *__restrict__ start, const int64_t *__restrict__ id,
                                   int32_t *__restrict__ sorted) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t v = items[k];
  const int32_t c = cell_of[v];
  const int64_t b = start[c], e = start[c + 1];
  const int64_t key = id ? id[v] : v;  // without ids: the index
  int64_t r = 0;
  for (int64_t m = b; m < e; ++m) r += (id ? id[items[m]] : items[m]) < key ? 1 : 0;
  sorted[b + r] = v;
}

// Cell of each particle in upload order (inspection).
__global__ void k_cell_of_particle(int64_t nslots, const int32_t *__restrict__ orig, const int32_t *__restrict__ cell,
                                   int64_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  const int32_t o = orig[i];
  if (o >= 0) out[o] = cell[i];
}

__global__ void k_gather_cell(int64_t n, const int32_t *__restrict__ src, const int32_t *__restrict__ cell_old,
                              int32_t *__restrict__ cell_new) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) cell_new[s] = cell_old[src[s]];
}

CellGeom make_cell_geom(const double lo[3], const double hi[3], double ilen, int64_t *ncells) {
  // CellBlock3D::rebuild with cellSizeFactor 1 (CellBlock3D.h:359-405)
  CellGeom g{};
  g.cpil = 1;
  int64_t nc = 1;
  for (int d = 0; d < 3; ++d) {
    g.box_min[d] = lo[d];
    g.box_max[d] = hi[d];
    const double L = hi[d] - lo[d];
    g.L[d] = L;
    long cpd = (long)std::floor(L / (ilen * 1.0));
    if (cpd < 1) cpd = 1;
    g.nh[d] = (int32_t)(cpd + 2 * g.cpil);
    g.cell_len[d] = L / (double)cpd;
    g.recip[d] = (double)cpd / L;
    nc *= g.nh[d];
  }
  *ncells = nc;
  return g;
}

void lc_bin(Ctx &c) {
  int64_t ncells = 0;
  c.cg = make_cell_geom(c.prm.box_min, c.prm.box_max, c.prm.cutoff + c.prm.skin, &ncells);
  CellGeom &g = c.cg;
  c.ncells = ncells;
  const int64_t ns = c.nslots;
  c.cell_of.ensure(ns);
  c.rank.ensure(ns);
  c.perm.ensure(ns);
  c.cell_count.ensure(ncells + 1);
  c.cell_start.ensure(ncells + 1);
  StateView s = view(c.st[c.cur]);
  LJG_HIP(hipMemsetAsync(c.cell_count.p, 0, (ncells + 1) * sizeof(int32_t), c.stream));
  hipLaunchKernelGGL(k_cell_hist, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, s.x, s.y, s.z, s.orig, g,
                     c.cell_of.p, c.rank.p, c.cell_count.p);
  LJG_HIP(hipGetLastError());
  exclusive_scan_i32_to_i64(c, c.cell_count.p, c.cell_start.p, ncells);
  hipLaunchKernelGGL(k_cell_scatter, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, c.cell_of.p, c.rank.p,
                     c.cell_start.p, c.perm.p);
  LJG_HIP(hipGetLastError());
  c.cell_sorted.ensure((size_t)c.n);
  hipLaunchKernelGGL(k_cell_order_by_id, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.p,
                     c.cell_of.p, c.cell_start.p, s.id, c.cell_sorted.p);
  LJG_HIP(hipGetLastError());
  std::swap(c.perm, c.cell_sorted);
  c.cell_of_particle.ensure(c.n);
  hipLaunchKernelGGL(k_cell_of_particle, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, s.orig, c.cell_of.p,
                     c.cell_of_particle.p);
  LJG_HIP(hipGetLastError());
  // Gather the state into cell order: the real particles fill slots [0, n) (padding slots are dropped).
  permute_state(c, c.perm.p, c.n);
  // cell of every new slot, for the force kernel (old slot's cell, gathered through the same map)
  hipLaunchKernelGGL(k_gather_cell, dim3(blocks_for(c.n, 256)), dim3(256), 0, c.stream, c.n, c.perm.p, c.cell_of.p,
                     c.rank.p);
  LJG_HIP(hipGetLastError());
  std::swap(c.cell_of, c.rank);
  c.layout = Layout::Cells;
}

void cell_bin_launch(hipStream_t st, int64_t n, const double *x, const double *y, const double *z, const CellGeom &g,
                     int32_t *cell_of, int32_t *rank, int32_t *count, const int64_t *start, int32_t *items,
                     const int64_t *id, int32_t *sorted, int phase) {
  if (phase == 0) {
    hipLaunchKernelGGL(k_cell_hist, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, x, y, z,
                       (const int32_t *)nullptr, g, cell_of, rank, count);
  } else {
    hipLaunchKernelGGL(k_cell_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, cell_of, rank, start, items);
    hipLaunchKernelGGL(k_cell_order_by_id, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, items, cell_of, start, id,
                       sorted);
  }
  LJG_HIP(hipGetLastError());
}

}  // namespace ljg
