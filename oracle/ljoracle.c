/*
 * ljoracle.c -- CPU restatement (FP64) of AutoPas's LJ short-range path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, never part of the product path.
 * Compiled with -O2 -ffp-contract=off -fopenmp (see oracle/Makefile).  Each function cites the
 * reference file:line it restates; paths are relative to the AutoPas source tree.
 */
#include "ljoracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MAX3(a, b) ((a) > (b) ? (a) : (b))
#define MIN3(a, b) ((a) < (b) ? (a) : (b))

/* ------------------------------------------------------------------------------------------------ */
/* LJ pair term.  LJFunctor.h:143-158 (AoS) and :288-320 (SoA):
 *   dr2 = dot(dr,dr); skip if dr2 > rc2; invdr2 = 1/dr2; lj6 = (sig2*invdr2)^3; lj12 = lj6^2;
 *   lj12m6 = lj12 - lj6; fac = eps24*(lj12 + lj12m6)*invdr2; f = dr*fac; pot6 = eps24*lj12m6 + shift6. */
int ljo_lj_pair(const double dr[3], double eps24, double sig2, double shift6, double rc2, double f[3],
                double *pot6) {
  const double dr2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
  if (dr2 > rc2) {
    f[0] = f[1] = f[2] = 0.;
    if (pot6) *pot6 = 0.;
    return 0;
  }
  const double invdr2 = 1. / dr2;
  double lj6 = sig2 * invdr2;
  lj6 = lj6 * lj6 * lj6;
  const double lj12 = lj6 * lj6;
  const double lj12m6 = lj12 - lj6;
  const double fac = eps24 * (lj12 + lj12m6) * invdr2;
  f[0] = dr[0] * fac;
  f[1] = dr[1] * fac;
  f[2] = dr[2] * fac;
  if (pot6) *pot6 = eps24 * lj12m6 + shift6;
  return 1;
}

/* ParticlePropertiesLibrary.h:575-582 */
double ljo_calc_shift6(double eps24, double sig2, double rc2) {
  const double s2 = sig2 / rc2;
  const double s6 = s2 * s2 * s2;
  return eps24 * (s6 - s6 * s6);
}

/* ParticlePropertiesLibrary.h:443-473: eps24 = 24*sqrt(e_i e_j); sigma = (s_i+s_j)/2; sig2 = sigma^2. */
void ljo_mixing(int ntypes, const double *eps, const double *sig, double cutoff, int apply_shift,
                double *eps24, double *sig2, double *shift6) {
  const double rc2 = cutoff * cutoff;
  for (int a = 0; a < ntypes; ++a) {
    for (int b = 0; b < ntypes; ++b) {
      const int k = ntypes * a + b;
      const double e24 = 24 * sqrt(eps[a] * eps[b]);
      const double s = (sig[a] + sig[b]) / 2.0;
      const double s2 = s * s;
      eps24[k] = e24;
      sig2[k] = s2;
      shift6[k] = apply_shift ? ljo_calc_shift6(e24, s2, rc2) : 0.;
    }
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* CellBlock3D::rebuild, CellBlock3D.h:359-405. */
void ljo_cellblock_build(const double box_min[3], const double box_max[3], double ilen, double csf,
                         ljo_cellblock *cb) {
  memset(cb, 0, sizeof(*cb));
  cb->cells_per_il = csf >= 1.0 ? 1 : (int64_t)ceil(1.0 / csf);
  int64_t num_cells = 1;
  for (int d = 0; d < 3; ++d) {
    cb->box_min[d] = box_min[d];
    cb->box_max[d] = box_max[d];
    const double box_len = box_max[d] - box_min[d];
    int64_t cpd = (int64_t)floor(box_len / (ilen * csf));
    if (cpd < 1) cpd = 1;
    cb->cells_per_dim_halo[d] = cpd + 2 * cb->cells_per_il;
    cb->cell_len[d] = box_len / (double)cpd;
    cb->recip[d] = (double)cpd / box_len;
    cb->halo_min[d] = box_min[d] - cb->cells_per_il * cb->cell_len[d];
    cb->halo_max[d] = box_max[d] + cb->cells_per_il * cb->cell_len[d];
    num_cells *= cb->cells_per_dim_halo[d];
  }
  const int64_t *n = cb->cells_per_dim_halo;
  cb->first_owned = n[0] * n[1] * cb->cells_per_il + n[0] * cb->cells_per_il + cb->cells_per_il;
  cb->last_owned = num_cells - 1 - cb->first_owned;
}

/* CellBlock3D::get3DIndexOfPosition, CellBlock3D.h:320-349. */
void ljo_cell_index3d(const ljo_cellblock *cb, const double pos[3], int64_t idx[3]) {
  for (int d = 0; d < 3; ++d) {
    const int64_t value = (int64_t)floor((pos[d] - cb->box_min[d]) * cb->recip[d]) + cb->cells_per_il;
    const int64_t nonneg = value > 0 ? value : 0;
    const int64_t upper = cb->cells_per_dim_halo[d] - 1;
    int64_t c = nonneg < upper ? nonneg : upper;
    if (pos[d] >= cb->box_max[d]) {
      const int64_t first_upper_halo = cb->cells_per_dim_halo[d] - cb->cells_per_il;
      c = c > first_upper_halo ? c : first_upper_halo;
    } else if (pos[d] < cb->box_min[d] && c == cb->cells_per_il) {
      --c;
    } else if (pos[d] < cb->box_max[d] && c == cb->cells_per_dim_halo[d] - cb->cells_per_il) {
      --c;
    }
    idx[d] = c;
  }
}

/* threeToOneD, ThreeDimensionalMapping.h:29-31: (z*ny + y)*nx + x. */
int64_t ljo_cell_index1d(const ljo_cellblock *cb, const double pos[3]) {
  int64_t c[3];
  ljo_cell_index3d(cb, pos, c);
  return (c[2] * cb->cells_per_dim_halo[1] + c[1]) * cb->cells_per_dim_halo[0] + c[0];
}

void ljo_cell_index_batch(const ljo_cellblock *cb, int64_t n, const double *x, const double *y, const double *z,
                          int64_t *cell) {
  for (int64_t i = 0; i < n; ++i) {
    const double p[3] = {x[i], y[i], z[i]};
    cell[i] = ljo_cell_index1d(cb, p);
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* computePairwiseCellOffsetsC08<noSorting>, LCC08CellHandlerUtility.cpp:66-158, with includeCellPair
 * :35-48 and toMaskXY :21-32.  Directions are visited in the order frontLeft, backLeft, frontRight,
 * backRight (LCC08CellHandlerUtility.h:120-122). */
static int c08_include(int dir, const int ov[3], int x, int y, int z) {
  switch (dir) {
    case 1: return y != ov[1] && z != 0;                 /* backLeft */
    case 2: return x != ov[0] && (y != 0 || z != 0);     /* frontRight */
    case 3: return y != ov[1] && x != ov[0] && z != 0;   /* backRight */
    default: return 1;                                   /* frontLeft */
  }
}

int ljo_c08_offsets(const int64_t cells_per_dim[3], const double cell_len[3], double ilen, int64_t *off2,
                    int64_t *off1, int cap) {
  int ov[3];
  for (int d = 0; d < 3; ++d) ov[d] = (int)ceil(ilen / cell_len[d]);
  const int64_t nx = cells_per_dim[0], ny = cells_per_dim[1];
  const double il2 = ilen * ilen;
  const int maskx[4] = {0, 0, 1, 1}, masky[4] = {0, 1, 0, 1};
  int count = 0;
  for (int x = 0; x <= ov[0]; ++x) {
    for (int y = 0; y <= ov[1]; ++y) {
      const int64_t o1 = (0 * ny + y) * nx + x;
      for (int z = 0; z <= ov[2]; ++z) {
        const double dz = MAX3(0.0, z - 1.0) * cell_len[2];
        for (int dir = 0; dir < 4; ++dir) {
          if (!c08_include(dir, ov, x, y, z)) continue;
          const int mx = maskx[dir], my = masky[dir];
          const int64_t o2 = ((int64_t)z * ny + my * ov[1]) * nx + mx * ov[0];
          const double dx = MAX3(0.0, (ov[0] - x) * mx + x * (1 - mx) - 1.0) * cell_len[0];
          const double dy = MAX3(0.0, (ov[1] - y) * my + y * (1 - my) - 1.0) * cell_len[1];
          if (dx * dx + dy * dy + dz * dz <= il2) {
            if (count < cap) {
              off2[count] = o2;
              off1[count] = o1;
            }
            ++count;
          }
        }
      }
    }
  }
  return count <= cap ? count : -count;
}

/* ------------------------------------------------------------------------------------------------ */
/* RegularGridDecomposition.cpp:563-587: a particle below boxMin goes to min(nextafter(max, min), p+L);
 * one at or above boxMax goes to max(min, p-L). */
void ljo_wrap_periodic(int64_t n, double *x, double *y, double *z, const double box_min[3],
                       const double box_max[3]) {
  double *c[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) {
    const double L = box_max[d] - box_min[d];
    const double just_inside = nextafter(box_max[d], box_min[d]);
    for (int64_t i = 0; i < n; ++i) {
      double p = c[d][i];
      if (p < box_min[d]) {
        const double q = p + L;
        p = q < just_inside ? q : just_inside;
      } else if (p >= box_max[d]) {
        const double q = p - L;
        p = q > box_min[d] ? q : box_min[d];
      }
      c[d][i] = p;
    }
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* Shared pair evaluation for the force oracles.  The type tables follow ParticlePropertiesLibrary's
 * row-major layout. */
typedef struct {
  const int32_t *type;
  int ntypes;
  const double *eps24, *sig2, *shift6;
  double rc2;
} pairctx;

static inline int eval_pair(const pairctx *pc, int64_t i, int64_t j, const double dr[3], double f[3],
                            double *pot6) {
  const int k = pc->type ? pc->type[i] * pc->ntypes + pc->type[j] : 0;
  return ljo_lj_pair(dr, pc->eps24[k], pc->sig2[k], pc->shift6[k], pc->rc2, f, pot6);
}

/* FP64 O(N^2) sum over periodic images (explicit image loop, so boxes shorter than 2*rc are handled).
 * Semantics of DirectSum + DSSequentialTraversal (DirectSum.h:66-114, DSSequentialTraversal.h:93-120)
 * with the periodic images standing in for md-flexible's halo copies.  Full shell: each unique pair
 * is evaluated from both sides, force accumulated on i only. */
int64_t ljo_direct_sum(int64_t n, const double *x, const double *y, const double *z, const int32_t *type,
                       int ntypes, const double *eps24, const double *sig2, const double *shift6, double cutoff,
                       const double *box_len, int threads, double *fx, double *fy, double *fz, double *epot,
                       double *virial) {
  pairctx pc = {type, ntypes, eps24, sig2, shift6, cutoff * cutoff};
  int K[3] = {0, 0, 0};
  int min_image = 0;
  if (box_len) {
    min_image = 1;
    for (int d = 0; d < 3; ++d) {
      K[d] = (int)ceil(cutoff / box_len[d]);
      if (!(box_len[d] > 2.0 * cutoff)) min_image = 0;
    }
  }
  double esum = 0., vsum = 0.;
  int64_t hits = 0;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : esum, vsum, hits)
  for (int64_t i = 0; i < n; ++i) {
    double ax = 0., ay = 0., az = 0.;
    for (int64_t j = 0; j < n; ++j) {
      if (min_image) {
        if (j == i) continue;
        double dr[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
        for (int d = 0; d < 3; ++d) {
          if (dr[d] > 0.5 * box_len[d]) dr[d] -= box_len[d];
          else if (dr[d] < -0.5 * box_len[d]) dr[d] += box_len[d];
        }
        double f[3], p6;
        if (eval_pair(&pc, i, j, dr, f, &p6)) {
          ax += f[0]; ay += f[1]; az += f[2];
          esum += p6;
          vsum += dr[0] * f[0] + dr[1] * f[1] + dr[2] * f[2];
          ++hits;
        }
        continue;
      }
      for (int kz = -K[2]; kz <= K[2]; ++kz)
        for (int ky = -K[1]; ky <= K[1]; ++ky)
          for (int kx = -K[0]; kx <= K[0]; ++kx) {
            if (j == i && kx == 0 && ky == 0 && kz == 0) continue;
            double sx = 0., sy = 0., sz = 0.;
            if (box_len) {
              sx = kx * box_len[0];
              sy = ky * box_len[1];
              sz = kz * box_len[2];
            }
            double dr[3] = {x[i] - (x[j] + sx), y[i] - (y[j] + sy), z[i] - (z[j] + sz)};
            double f[3], p6;
            if (eval_pair(&pc, i, j, dr, f, &p6)) {
              ax += f[0]; ay += f[1]; az += f[2];
              esum += p6;
              vsum += dr[0] * f[0] + dr[1] * f[1] + dr[2] * f[2];
              ++hits;
            }
          }
    }
    fx[i] = ax; fy[i] = ay; fz[i] = az;
  }
  /* Each unique pair was counted twice: LJFunctor::endTraversal halves and divides the energy by 6
   * (LJFunctor.h:661-685). */
  if (epot) *epot = esum * 0.5 / 6.;
  if (virial) *virial = vsum * 0.5;
  return hits / 2;
}

/* ------------------------------------------------------------------------------------------------ */
/* Counting sort of particles into the cells of a CellBlock3D (LinkedCells binning, LinkedCells.h:79-90). */
typedef struct {
  ljo_cellblock cb;
  int64_t ncells;
  int64_t *start; /* ncells+1 */
  int64_t *items; /* particle indices grouped by cell, input order inside a cell */
} cellbins;

static void bin_particles(cellbins *cbins, int64_t n, const double *x, const double *y, const double *z) {
  const ljo_cellblock *cb = &cbins->cb;
  cbins->ncells = cb->cells_per_dim_halo[0] * cb->cells_per_dim_halo[1] * cb->cells_per_dim_halo[2];
  cbins->start = calloc(cbins->ncells + 1, sizeof(int64_t));
  cbins->items = malloc((n > 0 ? n : 1) * sizeof(int64_t));
  int64_t *cell = malloc((n > 0 ? n : 1) * sizeof(int64_t));
  ljo_cell_index_batch(cb, n, x, y, z, cell);
  for (int64_t i = 0; i < n; ++i) cbins->start[cell[i] + 1]++;
  for (int64_t c = 0; c < cbins->ncells; ++c) cbins->start[c + 1] += cbins->start[c];
  int64_t *fill = malloc((cbins->ncells + 1) * sizeof(int64_t));
  memcpy(fill, cbins->start, (cbins->ncells + 1) * sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) cbins->items[fill[cell[i]]++] = i;
  free(fill);
  free(cell);
}

static void free_bins(cellbins *cbins) {
  free(cbins->start);
  free(cbins->items);
}

/* Linked cells, full shell, periodic: cells of CellBlock3D geometry (owned range only, positions must
 * lie inside the box), 27 neighbour cells with the periodic image shift of the wrapped cell.  With at
 * least one cell of side >= ilen per dimension, every image within rc is visited exactly once. */
int64_t ljo_lc_forces(int64_t n, const double *x, const double *y, const double *z, const int32_t *type,
                      int ntypes, const double *eps24, const double *sig2, const double *shift6, double cutoff,
                      double skin, const double box_min[3], const double box_max[3], int threads, double *fx,
                      double *fy, double *fz, double *epot, double *virial) {
  pairctx pc = {type, ntypes, eps24, sig2, shift6, cutoff * cutoff};
  cellbins bins;
  ljo_cellblock_build(box_min, box_max, cutoff + skin, 1.0, &bins.cb);
  bin_particles(&bins, n, x, y, z);
  const int64_t *nh = bins.cb.cells_per_dim_halo;
  const int64_t no[3] = {nh[0] - 2, nh[1] - 2, nh[2] - 2};
  double L[3];
  for (int d = 0; d < 3; ++d) L[d] = box_max[d] - box_min[d];
  double esum = 0., vsum = 0.;
  int64_t hits = 0;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : esum, vsum, hits)
  for (int64_t i = 0; i < n; ++i) {
    const double pi[3] = {x[i], y[i], z[i]};
    int64_t ci[3];
    ljo_cell_index3d(&bins.cb, pi, ci);
    double a[3] = {0., 0., 0.};
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
          const int o[3] = {ox, oy, oz};
          int64_t cj[3];
          double sh[3];
          for (int d = 0; d < 3; ++d) {
            int64_t c = ci[d] - 1 + o[d]; /* owned-range coordinate */
            sh[d] = 0.;
            if (c < 0) { c += no[d]; sh[d] = -L[d]; }
            else if (c >= no[d]) { c -= no[d]; sh[d] = L[d]; }
            cj[d] = c + 1;
          }
          const int64_t cidx = (cj[2] * nh[1] + cj[1]) * nh[0] + cj[0];
          for (int64_t s = bins.start[cidx]; s < bins.start[cidx + 1]; ++s) {
            const int64_t j = bins.items[s];
            if (j == i && sh[0] == 0. && sh[1] == 0. && sh[2] == 0.) continue;
            double dr[3] = {pi[0] - (x[j] + sh[0]), pi[1] - (y[j] + sh[1]), pi[2] - (z[j] + sh[2])};
            double f[3], p6;
            if (eval_pair(&pc, i, j, dr, f, &p6)) {
              a[0] += f[0]; a[1] += f[1]; a[2] += f[2];
              esum += p6;
              vsum += dr[0] * f[0] + dr[1] * f[1] + dr[2] * f[2];
              ++hits;
            }
          }
        }
    fx[i] = a[0]; fy[i] = a[1]; fz[i] = a[2];
  }
  free_bins(&bins);
  if (epot) *epot = esum * 0.5 / 6.;
  if (virial) *virial = vsum * 0.5;
  return hits / 2;
}

/* ------------------------------------------------------------------------------------------------ */
/* md-flexible's periodic halo copies, RegularGridDecomposition.cpp:159-236 with the single-rank
 * collectors :479-545: per dimension, owned particles in [min, min+ilen) get a +L copy and owned
 * particles in [max-ilen, max) a -L copy (region iterators select owned particles inside the box);
 * halos made for earlier dimensions in [min-skin, min+ilen) get a +L copy, else those in
 * [max-ilen, max+skin) a -L copy. */
typedef struct {
  int64_t n, cap;
  double *x, *y, *z;
  int64_t *src; /* owned particle each copy images */
} halobuf;

static void halo_push(halobuf *h, double x, double y, double z, int64_t src) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 1024;
    h->x = realloc(h->x, h->cap * sizeof(double));
    h->y = realloc(h->y, h->cap * sizeof(double));
    h->z = realloc(h->z, h->cap * sizeof(double));
    h->src = realloc(h->src, h->cap * sizeof(int64_t));
  }
  h->x[h->n] = x; h->y[h->n] = y; h->z[h->n] = z; h->src[h->n] = src;
  h->n++;
}

static void make_halos(int64_t n, const double *x, const double *y, const double *z, const double bmin[3],
                       const double bmax[3], double cutoff, double skin, halobuf *h) {
  const double ilen = cutoff + skin;
  memset(h, 0, sizeof(*h));
  for (int d = 0; d < 3; ++d) {
    const double L = bmax[d] - bmin[d];
    const int64_t nhalo_before = h->n;
    /* owned particles */
    for (int64_t i = 0; i < n; ++i) {
      const double p[3] = {x[i], y[i], z[i]};
      int inside = 1;
      for (int e = 0; e < 3; ++e) inside &= (p[e] >= bmin[e] && p[e] < bmax[e]);
      if (!inside) continue;
      if (p[d] < bmin[d] + ilen) {
        double q[3] = {p[0], p[1], p[2]};
        q[d] = q[d] + L;
        halo_push(h, q[0], q[1], q[2], i);
      }
      if (p[d] >= bmax[d] - ilen) {
        double q[3] = {p[0], p[1], p[2]};
        q[d] = q[d] + (-L);
        halo_push(h, q[0], q[1], q[2], i);
      }
    }
    /* halos from earlier dimensions */
    const double lmin = bmin[d] - skin, lmax = bmin[d] + cutoff + skin;
    const double rmin = bmax[d] - cutoff - skin, rmax = bmax[d] + skin;
    for (int64_t k = 0; k < nhalo_before; ++k) {
      double q[3] = {h->x[k], h->y[k], h->z[k]};
      if (q[d] >= lmin && q[d] < lmax) {
        q[d] = q[d] + L;
        halo_push(h, q[0], q[1], q[2], h->src[k]);
      } else if (q[d] >= rmin && q[d] < rmax) {
        q[d] = q[d] - L;
        halo_push(h, q[0], q[1], q[2], h->src[k]);
      }
    }
  }
}

/* SoA functor semantics for a cell pair / single cell, LJFunctor.h:204-364 (single, always N3 inside
 * the cell, j > i) and :386-559 (pair, N3).  Owned weight per particle for the globals. */
typedef struct {
  const double *x, *y, *z;
  const int32_t *type;
  const int8_t *owned;
  double *fx, *fy, *fz;
} soaall;

static void soa_single(const pairctx *pc, soaall *s, const int64_t *idx, int64_t cnt, double *e, double *v,
                       int64_t *hits) {
  for (int64_t a = 0; a < cnt; ++a) {
    const int64_t i = idx[a];
    double ax = 0., ay = 0., az = 0.;
    for (int64_t b = a + 1; b < cnt; ++b) {
      const int64_t j = idx[b];
      double dr[3] = {s->x[i] - s->x[j], s->y[i] - s->y[j], s->z[i] - s->z[j]};
      double f[3], p6;
      if (!eval_pair(pc, i, j, dr, f, &p6)) continue;
      ax += f[0]; ay += f[1]; az += f[2];
      s->fx[j] -= f[0]; s->fy[j] -= f[1]; s->fz[j] -= f[2];
      const double w = (double)(s->owned[i]) + (double)(s->owned[j]);
      *e += p6 * w;
      *v += (dr[0] * f[0] + dr[1] * f[1] + dr[2] * f[2]) * w;
      *hits += (int64_t)(s->owned[i] + s->owned[j]);
    }
    s->fx[i] += ax; s->fy[i] += ay; s->fz[i] += az;
  }
}

static void soa_pair_n3(const pairctx *pc, soaall *s, const int64_t *i1, int64_t c1, const int64_t *i2, int64_t c2,
                        double *e, double *v, int64_t *hits) {
  for (int64_t a = 0; a < c1; ++a) {
    const int64_t i = i1[a];
    double ax = 0., ay = 0., az = 0.;
    for (int64_t b = 0; b < c2; ++b) {
      const int64_t j = i2[b];
      double dr[3] = {s->x[i] - s->x[j], s->y[i] - s->y[j], s->z[i] - s->z[j]};
      double f[3], p6;
      if (!eval_pair(pc, i, j, dr, f, &p6)) continue;
      ax += f[0]; ay += f[1]; az += f[2];
      s->fx[j] -= f[0]; s->fy[j] -= f[1]; s->fz[j] -= f[2];
      const double w = (double)(s->owned[i]) + (double)(s->owned[j]);
      *e += p6 * w;
      *v += (dr[0] * f[0] + dr[1] * f[1] + dr[2] * f[2]) * w;
      /* a unique periodic pair appears once as owned-owned, or twice as owned-halo (once from each
       * side); count owned-owned as 2 halves and owned-halo as 1 half so the total is exact. */
      *hits += (int64_t)(s->owned[i] + s->owned[j]);
    }
    s->fx[i] += ax; s->fy[i] += ay; s->fz[i] += az;
  }
}

int64_t ljo_c08_forces_halo(int64_t n, const double *x, const double *y, const double *z, const int32_t *type,
                            int ntypes, const double *eps24, const double *sig2, const double *shift6,
                            double cutoff, double skin, const double box_min[3], const double box_max[3],
                            int threads, int repeats, double *fx, double *fy, double *fz, double *epot,
                            double *virial, double *traverse_seconds) {
  halobuf h;
  make_halos(n, x, y, z, box_min, box_max, cutoff, skin, &h);
  const int64_t ntot = n + h.n;
  double *ax = malloc(ntot * sizeof(double)), *ay = malloc(ntot * sizeof(double)),
         *az = malloc(ntot * sizeof(double));
  double *afx = calloc(ntot, sizeof(double)), *afy = calloc(ntot, sizeof(double)),
         *afz = calloc(ntot, sizeof(double));
  int32_t *at = malloc(ntot * sizeof(int32_t));
  int8_t *own = malloc(ntot);
  for (int64_t i = 0; i < n; ++i) {
    ax[i] = x[i]; ay[i] = y[i]; az[i] = z[i];
    at[i] = type ? type[i] : 0;
    own[i] = 1;
  }
  for (int64_t k = 0; k < h.n; ++k) {
    ax[n + k] = h.x[k]; ay[n + k] = h.y[k]; az[n + k] = h.z[k];
    at[n + k] = type ? type[h.src[k]] : 0;
    own[n + k] = 0;
  }
  pairctx pc = {at, type ? ntypes : 1, eps24, sig2, shift6, cutoff * cutoff};
  cellbins bins;
  ljo_cellblock_build(box_min, box_max, cutoff + skin, 1.0, &bins.cb);
  bin_particles(&bins, ntot, ax, ay, az);
  const int64_t *nh = bins.cb.cells_per_dim_halo;
  int64_t off2[64], off1[64];
  const int noff = ljo_c08_offsets(nh, bins.cb.cell_len, cutoff + skin, off2, off1, 64);
  soaall s = {ax, ay, az, at, own, afx, afy, afz};
  double esum = 0., vsum = 0.;
  int64_t halfhits = 0;
  if (threads > 0) omp_set_num_threads(threads);
  double t0 = omp_get_wtime();
  if (repeats < 1) repeats = 1;
  for (int rep = 0; rep < repeats; ++rep) {
    esum = 0.; vsum = 0.; halfhits = 0;
    memset(afx, 0, ntot * sizeof(double));
    memset(afy, 0, ntot * sizeof(double));
    memset(afz, 0, ntot * sizeof(double));
    /* ColorBasedTraversal.h:130-175: 8 colours, stride 2, base cells [0, n-overlap) per dimension. */
    for (int col = 0; col < 8; ++col) {
      const int sx = col % 2, sy = (col / 2) % 2, sz = col / 4;
#pragma omp parallel for schedule(dynamic, 1) collapse(3) reduction(+ : esum, vsum, halfhits)
      for (int64_t cz = sz; cz < nh[2] - 1; cz += 2)
        for (int64_t cy = sy; cy < nh[1] - 1; cy += 2)
          for (int64_t cx = sx; cx < nh[0] - 1; cx += 2) {
            const int64_t base = (cz * nh[1] + cy) * nh[0] + cx;
            for (int k = 0; k < noff; ++k) {
              /* LCC08CellHandler.h:154-170: (first, second) = (base + off2, base + off1). */
              const int64_t c1 = base + off2[k], c2 = base + off1[k];
              const int64_t s1 = bins.start[c1], e1 = bins.start[c1 + 1];
              const int64_t s2 = bins.start[c2], e2 = bins.start[c2 + 1];
              int64_t c1xyz[3] = {c1 % nh[0], (c1 / nh[0]) % nh[1], c1 / (nh[0] * nh[1])};
              int64_t c2xyz[3] = {c2 % nh[0], (c2 / nh[0]) % nh[1], c2 / (nh[0] * nh[1])};
              int own1 = 1, own2 = 1;
              for (int d = 0; d < 3; ++d) {
                own1 &= (c1xyz[d] >= 1 && c1xyz[d] < nh[d] - 1);
                own2 &= (c2xyz[d] >= 1 && c2xyz[d] < nh[d] - 1);
              }
              if (c1 == c2) {
                /* CellFunctor::processCell: skip empty and halo-only cells. */
                if (e1 - s1 == 0 || !own1) continue;
                soa_single(&pc, &s, bins.items + s1, e1 - s1, &esum, &vsum, &halfhits);
              } else {
                /* CellFunctor::processCellPair: skip empty, and pairs where neither cell can hold owned. */
                if (e1 - s1 == 0 || e2 - s2 == 0) continue;
                if (!own1 && !own2) continue;
                soa_pair_n3(&pc, &s, bins.items + s1, e1 - s1, bins.items + s2, e2 - s2, &esum, &vsum,
                            &halfhits);
              }
            }
          }
    }
  }
  double t1 = omp_get_wtime();
  if (traverse_seconds) *traverse_seconds = (t1 - t0) / repeats;
  for (int64_t i = 0; i < n; ++i) {
    fx[i] = afx[i]; fy[i] = afy[i]; fz[i] = afz[i];
  }
  if (epot) *epot = esum * 0.5 / 6.;
  if (virial) *virial = vsum * 0.5;
  free_bins(&bins);
  free(ax); free(ay); free(az); free(afx); free(afy); free(afz); free(at); free(own);
  free(h.x); free(h.y); free(h.z); free(h.src);
  return halfhits / 2;
}

/* ------------------------------------------------------------------------------------------------ */
/* VerletClusterLists, periodic variant.
 * Tower grid: ClusterTowerBlock2D::estimateOptimalGridSideLength (ClusterTowerBlock2D.h:140-168) with
 * density = N/V, side0 = cbrt(M/density), towers = ceil(L/side0), side = L/towers; no halo towers
 * because periodic images are addressed by shift vectors instead of halo copies.  Towers per
 * interaction length = max over x,y of ceil(ilen/side) (:95-101).  Tower of a position:
 * floor((p-min)*recip) clamped to [0, towers-1] (:220-245).  1D tower index x + y*towersX.
 * Clusters: tower sorted by z (ClusterTower.h:93, ties broken by particle id here), cut into groups of
 * M, the last one padded with dummies (:95-103).  Bounding box: xy min/max over the real particles,
 * z from the first and last real particle (Cluster.h:135-149).
 * Lists (full, i.e. the reference's newton3 == false form, VerletClusterListsRebuilder.h:237-409):
 * for i-cluster in tower (tx,ty): neighbour towers (ox,oy) in [-R,R]^2 (oy outer, ox inner) whose
 * xy distance max(0,|o|-1)*side (:289-301) squared-sum is <= ilen^2; the neighbour tower is the
 * periodic wrap of (tx+ox, ty+oy) and carries shift floor((t+o)/towers)*L; z images kz = -1,0,1;
 * every j-cluster whose shifted bounding box has boxDistanceSquared <= ilen^2 (ArrayMath.h:697-707) is
 * listed with its shift id (sx+1) + 3(sy+1) + 9(sz+1).  The entry (i, i, shift 0) is kept: it stands
 * for the reference's intra-cluster SoAFunctorSingle call (VCLClusterFunctor.h:56-72). */
struct ljo_vcl {
  int64_t num_slots, num_clusters, num_entries;
  int64_t towers[2], R;
  double side[2];
  int64_t *slot_particle;
  int32_t *cluster_tower;
  int64_t *list_offsets;
  int32_t *list_j, *list_shift;
  double *bbox;
};

typedef struct {
  double z;
  int64_t id;
  int64_t idx;
} zkey;

static int cmp_zkey(const void *a, const void *b) {
  const zkey *p = a, *q = b;
  if (p->z < q->z) return -1;
  if (p->z > q->z) return 1;
  if (p->id < q->id) return -1;
  if (p->id > q->id) return 1;
  return 0;
}

static double box_dist2(const double *a, const double *b, const double sh[3]) {
  /* boxDistanceSquared(aMin,aMax,bMin,bMax) with b shifted: sum of max(0,aMin-bMax)^2 + max(0,bMin-aMax)^2 */
  double s = 0.;
  for (int d = 0; d < 3; ++d) {
    const double bmin = b[d] + sh[d], bmax = b[3 + d] + sh[d];
    const double ab = MAX3(0.0, a[d] - bmax);
    const double ba = MAX3(0.0, bmin - a[3 + d]);
    s += ab * ab;
    s += ba * ba;
  }
  return s;
}

ljo_vcl *ljo_vcl_build_periodic(int64_t n, const double *x, const double *y, const double *z, const int64_t *id,
                                const double box_min[3], const double box_max[3], double ilen, int M) {
  ljo_vcl *v = calloc(1, sizeof(ljo_vcl));
  double L[3];
  for (int d = 0; d < 3; ++d) L[d] = box_max[d] - box_min[d];
  const double volume = L[0] * L[1] * L[2];
  const double density = (double)n / volume;
  const double side0 = cbrt((double)M / density);
  for (int d = 0; d < 2; ++d) {
    int64_t t = (int64_t)ceil(L[d] / side0);
    if (t < 1) t = 1;
    v->towers[d] = t;
    v->side[d] = L[d] / (double)t;
  }
  const int64_t R0 = (int64_t)ceil(ilen / v->side[0]), R1 = (int64_t)ceil(ilen / v->side[1]);
  v->R = R0 > R1 ? R0 : R1;
  const int64_t ntow = v->towers[0] * v->towers[1];
  /* tower of each particle */
  int64_t *tow = malloc((n ? n : 1) * sizeof(int64_t));
  int64_t *cnt = calloc(ntow + 1, sizeof(int64_t));
  const double recip[2] = {1. / v->side[0], 1. / v->side[1]};
  for (int64_t i = 0; i < n; ++i) {
    const double p[2] = {x[i], y[i]};
    int64_t t2[2];
    for (int d = 0; d < 2; ++d) {
      int64_t t = (int64_t)floor((p[d] - box_min[d]) * recip[d]);
      if (t < 0) t = 0;
      if (t > v->towers[d] - 1) t = v->towers[d] - 1;
      t2[d] = t;
    }
    tow[i] = t2[0] + t2[1] * v->towers[0];
    cnt[tow[i] + 1]++;
  }
  for (int64_t t = 0; t < ntow; ++t) cnt[t + 1] += cnt[t];
  zkey *keys = malloc((n ? n : 1) * sizeof(zkey));
  int64_t *fill = malloc((ntow + 1) * sizeof(int64_t));
  memcpy(fill, cnt, (ntow + 1) * sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) {
    zkey k = {z[i], id ? id[i] : i, i};
    keys[fill[tow[i]]++] = k;
  }
  /* clusters per tower, slots */
  int64_t *cstart = calloc(ntow + 1, sizeof(int64_t));
  for (int64_t t = 0; t < ntow; ++t) {
    const int64_t nt = cnt[t + 1] - cnt[t];
    qsort(keys + cnt[t], nt, sizeof(zkey), cmp_zkey);
    cstart[t + 1] = cstart[t] + (nt + M - 1) / M;
  }
  v->num_clusters = cstart[ntow];
  v->num_slots = v->num_clusters * M;
  v->slot_particle = malloc((v->num_slots ? v->num_slots : 1) * sizeof(int64_t));
  v->cluster_tower = malloc((v->num_clusters ? v->num_clusters : 1) * sizeof(int32_t));
  v->bbox = malloc((v->num_clusters ? v->num_clusters : 1) * 6 * sizeof(double));
  for (int64_t t = 0; t < ntow; ++t) {
    const int64_t nt = cnt[t + 1] - cnt[t];
    for (int64_t c = cstart[t]; c < cstart[t + 1]; ++c) {
      v->cluster_tower[c] = (int32_t)t;
      double *bb = v->bbox + 6 * c;
      const int64_t k0 = (c - cstart[t]) * M;
      int first = 1;
      for (int m = 0; m < M; ++m) {
        const int64_t k = k0 + m;
        if (k < nt) {
          const int64_t p = keys[cnt[t] + k].idx;
          v->slot_particle[c * M + m] = p;
          if (first) {
            bb[0] = bb[3] = x[p];
            bb[1] = bb[4] = y[p];
            bb[2] = z[p];
            first = 0;
          }
          bb[0] = MIN3(bb[0], x[p]); bb[3] = MAX3(bb[3], x[p]);
          bb[1] = MIN3(bb[1], y[p]); bb[4] = MAX3(bb[4], y[p]);
          bb[5] = z[p];
        } else {
          v->slot_particle[c * M + m] = -1;
        }
      }
    }
  }
  /* lists: count then fill, same loop order */
  v->list_offsets = calloc(v->num_clusters + 1, sizeof(int64_t));
  const double il2 = ilen * ilen;
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      for (int64_t c = 0; c < v->num_clusters; ++c) v->list_offsets[c + 1] += v->list_offsets[c];
      v->num_entries = v->list_offsets[v->num_clusters];
      v->list_j = malloc((v->num_entries ? v->num_entries : 1) * sizeof(int32_t));
      v->list_shift = malloc((v->num_entries ? v->num_entries : 1) * sizeof(int32_t));
    }
    for (int64_t ci = 0; ci < v->num_clusters; ++ci) {
      const int64_t t = v->cluster_tower[ci];
      const int64_t tx = t % v->towers[0], ty = t / v->towers[0];
      int64_t w = pass == 1 ? v->list_offsets[ci] : 0;
      int64_t count = 0;
      for (int64_t oy = -v->R; oy <= v->R; ++oy) {
        const double dty = (double)MAX3((int64_t)0, llabs(oy) - 1) * v->side[1];
        for (int64_t ox = -v->R; ox <= v->R; ++ox) {
          const double dtx = (double)MAX3((int64_t)0, llabs(ox) - 1) * v->side[0];
          if (dtx * dtx + dty * dty > il2) continue;
          int64_t jx = tx + ox, jy = ty + oy;
          int sx = 0, sy = 0;
          while (jx < 0) { jx += v->towers[0]; --sx; }
          while (jx >= v->towers[0]) { jx -= v->towers[0]; ++sx; }
          while (jy < 0) { jy += v->towers[1]; --sy; }
          while (jy >= v->towers[1]) { jy -= v->towers[1]; ++sy; }
          const int64_t tj = jx + jy * v->towers[0];
          for (int sz = -1; sz <= 1; ++sz) {
            const double sh[3] = {sx * L[0], sy * L[1], sz * L[2]};
            for (int64_t cj = cstart[tj]; cj < cstart[tj + 1]; ++cj) {
              if (box_dist2(v->bbox + 6 * ci, v->bbox + 6 * cj, sh) <= il2) {
                if (pass == 1) {
                  v->list_j[w] = (int32_t)cj;
                  v->list_shift[w] = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
                  ++w;
                }
                ++count;
              }
            }
          }
        }
      }
      if (pass == 0) v->list_offsets[ci + 1] = count;
    }
  }
  free(tow); free(cnt); free(keys); free(fill); free(cstart);
  return v;
}

void ljo_vcl_sizes(const ljo_vcl *v, int64_t *num_slots, int64_t *num_clusters, int64_t *num_entries,
                   int64_t towers_per_dim[2], double side[2], int64_t *towers_per_il) {
  *num_slots = v->num_slots;
  *num_clusters = v->num_clusters;
  *num_entries = v->num_entries;
  towers_per_dim[0] = v->towers[0];
  towers_per_dim[1] = v->towers[1];
  side[0] = v->side[0];
  side[1] = v->side[1];
  *towers_per_il = v->R;
}

void ljo_vcl_get(const ljo_vcl *v, int64_t *slot_particle, int32_t *cluster_tower, int64_t *list_offsets,
                 int32_t *list_j, int32_t *list_shift, double *bbox) {
  if (slot_particle) memcpy(slot_particle, v->slot_particle, v->num_slots * sizeof(int64_t));
  if (cluster_tower) memcpy(cluster_tower, v->cluster_tower, v->num_clusters * sizeof(int32_t));
  if (list_offsets) memcpy(list_offsets, v->list_offsets, (v->num_clusters + 1) * sizeof(int64_t));
  if (list_j) memcpy(list_j, v->list_j, v->num_entries * sizeof(int32_t));
  if (list_shift) memcpy(list_shift, v->list_shift, v->num_entries * sizeof(int32_t));
  if (bbox) memcpy(bbox, v->bbox, v->num_clusters * 6 * sizeof(double));
}

/* Newton-3 lists: only forward neighbours (VerletClusterListsRebuilder.h:278-306,339-354,366-409), so that
 * an interacting cluster pair is listed once instead of in both clusters' lists.  The reference orders
 * towers by interaction cell and then tower index, and within a tower takes the clusters after A; in the
 * periodic variant (no halo towers, neighbours reached through image shifts) the order must be antisymmetric
 * under (A, B, s) <-> (B, A, -s), so an entry (J, s) of cluster I is forward when J > I (clusters are
 * numbered tower by tower, z-sorted inside a tower: the reference's same-tower rule), or J == I and
 * s >= 13 (shift ids s and 26 - s are opposite images; 13 is no shift, the cluster itself, whose member
 * pairs a Newton-3 kernel takes once, as SoAFunctorSingle does, VCLClusterFunctor.h:56-72).  The order of
 * the kept entries is that of the full list. */
void ljo_vcl_forward_only(ljo_vcl *v) {
  int64_t w = 0, prev = 0;
  for (int64_t ci = 0; ci < v->num_clusters; ++ci) {
    const int64_t b = prev, e = v->list_offsets[ci + 1];
    prev = e;
    v->list_offsets[ci] = w;
    for (int64_t k = b; k < e; ++k) {
      const int64_t cj = v->list_j[k];
      const int sid = v->list_shift[k];
      if (cj > ci || (cj == ci && sid >= 13)) {
        v->list_j[w] = v->list_j[k];
        v->list_shift[w] = v->list_shift[k];
        ++w;
      }
    }
  }
  v->list_offsets[v->num_clusters] = w;
  v->num_entries = w;
}

void ljo_vcl_free(ljo_vcl *v) {
  if (!v) return;
  free(v->slot_particle); free(v->cluster_tower); free(v->list_offsets);
  free(v->list_j); free(v->list_shift); free(v->bbox);
  free(v);
}

/* ------------------------------------------------------------------------------------------------ */
static uint32_t to_fixed(double x, double mn, double inv_scale) {
  return (uint32_t)(uint64_t)llrint((x - mn) * inv_scale);
}

int64_t ljo_verlet_pairs_fixed(int64_t n, const double *x, const double *y, const double *z, const double box_min[3],
                               const double box_max[3], double ilen, int64_t *off, int32_t *nb, int64_t cap) {
  double scale[3], inv[3];
  for (int d = 0; d < 3; ++d) {
    const double L = box_max[d] - box_min[d];
    scale[d] = L / 4294967296.0;
    inv[d] = 4294967296.0 / L;
  }
  uint32_t *X = malloc((n ? n : 1) * 3 * sizeof(uint32_t));
  for (int64_t i = 0; i < n; ++i) {
    X[3 * i] = to_fixed(x[i], box_min[0], inv[0]);
    X[3 * i + 1] = to_fixed(y[i], box_min[1], inv[1]);
    X[3 * i + 2] = to_fixed(z[i], box_min[2], inv[2]);
  }
  const double il2 = ilen * ilen;
  int64_t tot = 0;
  off[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = (double)(int32_t)(X[3 * i] - X[3 * j]) * scale[0];
      const double dy = (double)(int32_t)(X[3 * i + 1] - X[3 * j + 1]) * scale[1];
      const double dz = (double)(int32_t)(X[3 * i + 2] - X[3 * j + 2]) * scale[2];
      if (dx * dx + dy * dy + dz * dz <= il2) {
        if (tot < cap) nb[tot] = (int32_t)j;
        ++tot;
      }
    }
    off[i + 1] = tot;
  }
  free(X);
  return tot <= cap ? tot : -tot;
}

/* ------------------------------------------------------------------------------------------------ */
/* TimeDiscretization.cpp:17-69: oldF = F; F = 0 (no global force); x += v*dt + F*(dt*dt/(2m)). */
void ljo_verlet_positions(int64_t n, double *x, double *y, double *z, const double *vx, const double *vy,
                          const double *vz, double *fx, double *fy, double *fz, double *ofx, double *ofy,
                          double *ofz, const double *mass_of_type, const int32_t *type, double dt) {
  for (int64_t i = 0; i < n; ++i) {
    const double m = mass_of_type[type ? type[i] : 0];
    const double fac = dt * dt / (2 * m);
    const double dx = vx[i] * dt + fx[i] * fac;
    const double dy = vy[i] * dt + fy[i] * fac;
    const double dz = vz[i] * dt + fz[i] * fac;
    ofx[i] = fx[i]; ofy[i] = fy[i]; ofz[i] = fz[i];
    fx[i] = 0.; fy[i] = 0.; fz[i] = 0.;
    x[i] += dx; y[i] += dy; z[i] += dz;
  }
}

/* TimeDiscretization.cpp:152-165: v += (F + oldF) * (dt / (2m)). */
void ljo_verlet_velocities(int64_t n, double *vx, double *vy, double *vz, const double *fx, const double *fy,
                           const double *fz, const double *ofx, const double *ofy, const double *ofz,
                           const double *mass_of_type, const int32_t *type, double dt) {
  for (int64_t i = 0; i < n; ++i) {
    const double m = mass_of_type[type ? type[i] : 0];
    const double fac = dt / (2 * m);
    vx[i] += (fx[i] + ofx[i]) * fac;
    vy[i] += (fy[i] + ofy[i]) * fac;
    vz[i] += (fz[i] + ofz[i]) * fac;
  }
}
