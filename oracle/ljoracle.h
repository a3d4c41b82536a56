/*
 * ljoracle.h -- CPU restatement (FP64, plain C) of AutoPas's Lennard-Jones short-range path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in autopas_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
 * path never links, imports or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the AutoPas tree).
 * The translation unit is compiled with -ffp-contract=off so the FP64 operation order written here
 * is the order executed (this is what "bit-exact" binning and pair sets are defined against).
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   ljo_lj_pair          known answers of LJFunctorTestNoGlobals.h:27-28, ForceCalculationTest.cpp:43-86,
 *                        closed-form LJPotential.h, F(2^(1/6) sigma) = 0, energy/virial closed form.
 *   ljo_mixing           closed form + ParticlePropertiesLibraryTest.cpp:103-104.
 *   ljo_cellblock_*      CellBlock3DTest.cpp boundary expectations + closed-form mesh indexing.
 *   ljo_c08_offsets      LCC08CellHandlerUtilityTest.cpp:27-57 (exact 14 pairs and their order).
 *   ljo_direct_sum       brute force over explicit periodic images; lattice closed forms (shell counts,
 *                        zero net force on a perfect lattice, analytic lattice energy).
 *   ljo_lc_forces        vs ljo_direct_sum (different algorithm, same physics).
 *   ljo_c08_forces_halo  vs ljo_direct_sum (halo copies + C08 colouring + Newton-3).
 *   ljo_vcl_*            VerletClusterListsTest.cpp:258-309 tower geometry; superset property vs the
 *                        brute-force pair set (every pair within rc covered exactly once per direction).
 *   ljo_verlet_*         TimeDiscretizationTest.cpp closed-form updates.
 */
#ifndef LJORACLE_H
#define LJORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- LJ 12-6 pair term: LJFunctor.h:143-158 (AoS), :288-320 (SoA) ---------------------------- */
/* Returns 1 if dr2 <= rc2 (inclusive) and writes f = dr*fac and pot6 = eps24*lj12m6 + shift6. */
int ljo_lj_pair(const double dr[3], double eps24, double sig2, double shift6, double rc2, double f[3],
                double *pot6);

/* ---- Lorentz-Berthelot mixing: ParticlePropertiesLibrary.h:443-489, calcShift6 :575-582 ---------- */
/* Tables are ntypes*ntypes, row-major [first*ntypes + second]. shift6 is zeroed if apply_shift == 0. */
void ljo_mixing(int ntypes, const double *eps, const double *sig, double cutoff, int apply_shift,
                double *eps24, double *sig2, double *shift6);
double ljo_calc_shift6(double eps24, double sig2, double rc2);

/* ---- CellBlock3D geometry and position -> cell: CellBlock3D.h:320-425, ThreeDimensionalMapping.h:29-31 */
typedef struct {
  double box_min[3], box_max[3];
  double cell_len[3], recip[3];
  double halo_min[3], halo_max[3];
  int64_t cells_per_dim_halo[3];
  int64_t cells_per_il;
  int64_t first_owned, last_owned;
} ljo_cellblock;

void ljo_cellblock_build(const double box_min[3], const double box_max[3], double ilen, double csf,
                         ljo_cellblock *cb);
void ljo_cell_index3d(const ljo_cellblock *cb, const double pos[3], int64_t idx[3]);
int64_t ljo_cell_index1d(const ljo_cellblock *cb, const double pos[3]);
/* Batch form (SoA input), for parity tests. */
void ljo_cell_index_batch(const ljo_cellblock *cb, int64_t n, const double *x, const double *y, const double *z,
                          int64_t *cell);

/* ---- C08 base-cell pair offsets: LCC08CellHandlerUtility.cpp:35-48, :66-158 (noSorting mode) ---- */
/* Writes pairs (offset2, offset1) in the reference's emission order. Returns the number of pairs,
 * or -needed if cap is too small. */
int ljo_c08_offsets(const int64_t cells_per_dim[3], const double cell_len[3], double ilen, int64_t *off2,
                    int64_t *off1, int cap);

/* ---- periodic wrap of leaving particles: RegularGridDecomposition.cpp:563-587 ------------------ */
void ljo_wrap_periodic(int64_t n, double *x, double *y, double *z, const double box_min[3],
                       const double box_max[3]);

/* ---- forces: all FP64, full sums, unique-pair counting ------------------------------------------ */
/* Type tables as produced by ljo_mixing. box_len == NULL means an open (non-periodic) box.
 * Energy is reported as the reference's getPotentialEnergy(), i.e. sum over unique pairs of
 * (eps24*lj12m6 + shift6)/6 (LJFunctor.h:661-685); virial as getVirial() = sum of dr.f over unique
 * pairs, x+y+z (LJFunctor.h:709-720).  Return value: number of unique pairs with r^2 <= rc^2. */
int64_t ljo_direct_sum(int64_t n, const double *x, const double *y, const double *z, const int32_t *type,
                       int ntypes, const double *eps24, const double *sig2, const double *shift6, double cutoff,
                       const double *box_len, int threads, double *fx, double *fy, double *fz, double *epot,
                       double *virial);

/* Linked cells, full shell (27 neighbour cells, periodic images by cell shift), same outputs. */
int64_t ljo_lc_forces(int64_t n, const double *x, const double *y, const double *z, const int32_t *type,
                      int ntypes, const double *eps24, const double *sig2, const double *shift6, double cutoff,
                      double skin, const double box_min[3], const double box_max[3], int threads, double *fx,
                      double *fy, double *fz, double *epot, double *virial);

/* md-flexible's CPU path: periodic halo copies (RegularGridDecomposition.cpp:159-236, 479-509),
 * LinkedCells + CellBlock3D binning, LC C08 colour traversal with Newton-3 (LCC08Traversal.h:72-79,
 * ColorBasedTraversal.h:130-175, LCC08CellHandler.h:154-170, CellFunctor.h:141-191), SoA functor
 * semantics of LJFunctor.h:204-559.  Forces are written for the n owned particles only.
 * *traverse_seconds (may be NULL) receives the wall time of the colour traversal alone. */
int64_t ljo_c08_forces_halo(int64_t n, const double *x, const double *y, const double *z, const int32_t *type,
                            int ntypes, const double *eps24, const double *sig2, const double *shift6,
                            double cutoff, double skin, const double box_min[3], const double box_max[3],
                            int threads, int repeats, double *fx, double *fy, double *fz, double *epot,
                            double *virial, double *traverse_seconds);

/* ---- VerletClusterLists, periodic variant: ClusterTowerBlock2D.h:140-168,220-245;
 *      ClusterTower.h:83-143; VerletClusterListsRebuilder.h:237-409; Cluster.h:135-149;
 *      ArrayMath.h:697-707 ------------------------------------------------------------------------ */
typedef struct ljo_vcl ljo_vcl;
ljo_vcl *ljo_vcl_build_periodic(int64_t n, const double *x, const double *y, const double *z, const int64_t *id,
                                const double box_min[3], const double box_max[3], double ilen, int cluster_size);
void ljo_vcl_sizes(const ljo_vcl *v, int64_t *num_slots, int64_t *num_clusters, int64_t *num_entries,
                   int64_t towers_per_dim[2], double side[2], int64_t *towers_per_il);
/* slot_particle[num_slots] (-1 for dummy padding), cluster_tower[num_clusters],
 * list_offsets[num_clusters+1], list_j[num_entries], list_shift[num_entries],
 * bbox[6*num_clusters] (xlo,ylo,zlo,xhi,yhi,zhi). */
void ljo_vcl_get(const ljo_vcl *v, int64_t *slot_particle, int32_t *cluster_tower, int64_t *list_offsets,
                 int32_t *list_j, int32_t *list_shift, double *bbox);
/* Keep only the forward (Newton-3) entries of every list, in list order (VerletClusterListsRebuilder.h:339-354). */
void ljo_vcl_forward_only(ljo_vcl *v);
void ljo_vcl_free(ljo_vcl *v);

/* ---- Verlet lists (VerletLists / SoAFunctorVerlet neighbour sets, VerletLists.h) ----------------
 * Pairs within ilen (inclusive) in a periodic box, evaluated on 32-bit fixed-point coordinates
 * X = round((x - min) * 2^32 / L) mod 2^32 (the engine's representation): separation = wrapping int32
 * difference times L / 2^32, r^2 summed in FP64 as (dx*dx + dy*dy) + dz*dz.  Brute force O(N^2).
 * neighbors of particle i are written to nb[off[i] .. off[i+1]) in increasing j.  Returns the total,
 * or -total if cap is too small (off is always complete). */
int64_t ljo_verlet_pairs_fixed(int64_t n, const double *x, const double *y, const double *z, const double box_min[3],
                               const double box_max[3], double ilen, int64_t *off, int32_t *nb, int64_t cap);

/* ---- velocity Verlet: TimeDiscretization.cpp:17-69 (positions), :152-165 (velocities) ---------- */
void ljo_verlet_positions(int64_t n, double *x, double *y, double *z, const double *vx, const double *vy,
                          const double *vz, double *fx, double *fy, double *fz, double *ofx, double *ofy,
                          double *ofz, const double *mass_of_type, const int32_t *type, double dt);
void ljo_verlet_velocities(int64_t n, double *vx, double *vy, double *vz, const double *fx, const double *fy,
                           const double *fz, const double *ofx, const double *ofy, const double *ofz,
                           const double *mass_of_type, const int32_t *type, double dt);

#ifdef __cplusplus
}
#endif
#endif
