/*
 * ljgpu.h -- C ABI of the MI355X (gfx950) engine for AutoPas's Lennard-Jones short-range path.
 *
 * The engine owns device memory; host arrays passed in are borrowed for the duration of the call.
 * Every entry point returns an ljg_status; no C++ exception crosses this boundary.  One context is
 * one periodic box on one HIP stream and is not thread-safe; separate contexts are independent.
 *
 * In AutoPas the path this replaces is reached through C++ templates rather than an FFI:
 *   AutoPas<Particle_T>::computeInteractions(Functor*)    src/autopas/AutoPasImpl.h:116-135
 *     -> LogicHandler::computeInteractions                src/autopas/LogicHandler.h:1066-1143
 *       -> container->rebuildNeighborLists / computeInteractions(traversal)
 * The entry points below are the calls a C++ container + traversal plugged in under that API makes;
 * each one cites the reference interface it stands in for.
 */
#ifndef LJGPU_H
#define LJGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ljg_ctx ljg_ctx;
typedef int ljg_status;

#define LJG_OK 0
#define LJG_ERR_ARG 1         /* invalid argument (null pointer, bad size, box smaller than ilen ...) */
#define LJG_ERR_HIP 2         /* a HIP runtime call failed; ljg_last_error() has the text */
#define LJG_ERR_STATE 3       /* call out of order (e.g. force before upload) */
#define LJG_ERR_CAPACITY 4    /* a structural limit was exceeded (e.g. a tower longer than the sort tile) */
#define LJG_ERR_UNSUPPORTED 5 /* configuration not provided by this engine */
#define LJG_ERR_NODEVICE 6    /* no HIP device available */

/* Force kernels.  The plain FP64 kernel is the linked-cells full-shell kernel; the cluster kernels run
 * on the Verlet-cluster lists. */
#define LJG_KERNEL_LC_FP64 0      /* LinkedCells, 27-cell full shell, FP64 everything */
#define LJG_KERNEL_CLUSTER_FP64 1 /* cluster-pair tiles, FP64 pair term and sums */
#define LJG_KERNEL_CLUSTER_MIXED 2 /* cluster-pair tiles, FP32 pair term, FP64 per-particle sums */
#define LJG_KERNEL_VERLET_MIXED 3  /* Verlet lists on linked cells (cluster size 1), fixed-point positions,
                                      FP32 pair term, FP64 per-particle sums; boxes >= 3 (rc + skin) */
#define LJG_KERNEL_CLUSTER_JROWS_MIXED 4 /* Verlet-cluster i-clusters against j-particle rows (the union of
                                      the members' neighbour lists), fixed-point positions, FP32 pair term,
                                      FP64 per-particle sums; boxes > 2 (rc + skin) */
#define LJG_KERNEL_CLUSTER_FP64_N3 5 /* LJG_KERNEL_CLUSTER_FP64 on Newton-3 lists: forward neighbours only, each
                                      interacting cluster pair listed once (VerletClusterListsRebuilder.h:339-354),
                                      each pair's force added to both particles (LJFunctor.h:316-318,
                                      SoAFunctorPairImpl<newton3 = true> :386-559) through device atomics; single
                                      boxes; not bit-reproducible run to run (the order of the additions) */
#define LJG_KERNEL_CLUSTER_MIXED_N3 6 /* LJG_KERNEL_CLUSTER_MIXED on Newton-3 lists, as LJG_KERNEL_CLUSTER_FP64_N3 */
#define LJG_KERNEL_LAST 6
/* ljg_step only: tune the layout while stepping (AutoPas's tuning phases, AutoTuner.cpp:85-175) and use
 * the fastest; see ljg_params.tuning_samples / tuning_interval and ljg_get_tuning. */
#define LJG_KERNEL_AUTO 100
#define LJG_MAX_TUNE_CANDIDATES 16

typedef struct {
  double box_min[3];
  double box_max[3];
  double cutoff;          /* rc; pairs with r^2 <= rc^2 interact (LJFunctor.h:300) */
  double skin;            /* Verlet skin; interaction length ilen = rc + skin (LogicHandler.h:95) */
  int32_t cluster_size;   /* 1, 2, 4 or 8 (VerletClusterLists clusterSize) */
  int32_t rebuild_frequency; /* steps between list rebuilds (AutoPas rebuildFrequency) */
  int32_t dynamic_rebuild;   /* also rebuild once any particle moved >= skin/2 (LogicHandler.h:1000-1015) */
  int32_t apply_shift;    /* truncated-shifted potential energy (LJFunctor applyShift) */
  int32_t device;         /* HIP device ordinal */
  int32_t tile_groups;    /* i-clusters per wavefront in the FP32 cluster kernels; 0 = default per cluster size */
  int32_t build_groups;   /* i-clusters per wavefront in the j-row expansion pass; 0 = default per cluster size */
  int32_t tuning_samples; /* LJG_KERNEL_AUTO: steps per candidate in a tuning phase (AutoPas maxSamples);
                             0 = max(3, rebuild_frequency) */
  int32_t tuning_interval; /* LJG_KERNEL_AUTO: steps between tuning phases (AutoPas tuningInterval); 0 = tune
                              once */
  int32_t list_mode;      /* cluster-entry list pass: 0 = by grid (a group of lanes per i-cluster when towers
                           * hold few clusters, else a wavefront per tower), 1 = wavefront per tower, 2 = lane
                           * per i-cluster, 3 = a group of 32 lanes per i-cluster; all give the same lists */
  int32_t shared_rebuilds; /* ensemble: 0 = every box keeps the rebuild schedule it has alone (frequency and
                            * displacement trigger per box, results bit-identical to the box alone); 1 = one
                            * schedule for all boxes (a trigger in any box rebuilds every box; results equal to
                            * the box alone up to the FP32 rounding of the pair sums), fewer and larger rebuilds */
  int32_t graphs;          /* ljg_step: 1 = the launches of each chunk of steps between two list rebuilds are
                            * captured into a hipGraph (stream capture) and launched as one graph; the graph
                            * of the previous chunk is updated in place when the launch sequence matches
                            * (hipGraphExecUpdate), else instantiated anew.  Per-launch timing events are
                            * then off (ljg_run_stats.force_ms stays 0). */
  int32_t whole_cluster_rows; /* FP32 cluster kernel (LJG_KERNEL_CLUSTER_MIXED): 0 = each i-cluster streams
                               * only the j-cluster members within rc + skin of one of its members (a row of
                               * (entry, member) pairs filtered at every list build); 1 = every member of
                               * every listed j-cluster, as VCLClusterFunctor does (VCLClusterFunctor.h:38-89) */
  int32_t persistent;         /* FP32 cluster kernel in ljg_step: 1 = one persistent launch per step, whose
                               * wavefronts (a few per SIMD) take tiles from a work queue in device memory,
                               * each starting at its XCD's share; 0 = a wavefront per tile */
  int32_t row_headroom;       /* cluster list builds: spare row capacity kept for the next build, in percent
                               * of the longest row plus 8 entries; 0 = 25 %, -1 = none (the longest row's
                               * length only).  A row that outgrows its capacity makes the build (and a force
                               * launched on it) run again with the lengths read back. */
} ljg_params;

typedef struct {
  double potential_energy; /* LJFunctor::getPotentialEnergy semantics (LJFunctor.h:691-705) */
  double virial;           /* LJFunctor::getVirial: x+y+z of sum r.F over pairs (LJFunctor.h:709-720) */
  int64_t pair_count;      /* unique pairs with r^2 <= rc^2 (each unordered periodic pair once) */
  int64_t candidate_evals; /* pair evaluations the kernel issued (incl. masked-out ones) */
} ljg_globals;

typedef struct {
  int64_t steps;
  int64_t rebuilds;
  int64_t pair_interactions; /* sum over steps of the unique pairs within rc */
  double force_ms;           /* summed HIP-event time of the timed force launches (when timing is on: every
                                8th step's launch, so that the event markers stay off most steps) */
  int64_t force_launches;
  double rebuild_ms;         /* summed HIP-event time of rebuilds (bin/sort/list) (when timing is on) */
  double wall_ms;            /* host wall time of the call */
} ljg_run_stats;

/* Result of the layout tuning (LJG_KERNEL_AUTO): the candidates, their scores (ms per step: rebuild time
 * / rebuild_frequency + the rest of the step, as AutoTuner::estimateRuntimeFromSamples with the mean
 * selector; -1 if not measured, e.g. the box is too small for that kernel) and the chosen one. */
typedef struct {
  int32_t candidates;
  int32_t best;    /* index of the chosen candidate, -1 before the first tuning phase ends */
  int32_t phases;  /* tuning phases started */
  int32_t tuning;  /* 1 while a tuning phase is in progress */
  int32_t kernel[LJG_MAX_TUNE_CANDIDATES];
  int32_t cluster_size[LJG_MAX_TUNE_CANDIDATES];
  int32_t tile_groups[LJG_MAX_TUNE_CANDIDATES];
  int32_t build_groups[LJG_MAX_TUNE_CANDIDATES];
  double effective_ms[LJG_MAX_TUNE_CANDIDATES];
} ljg_tuning;

/* Create/destroy a context.  Replaces AutoPas::init() for the container this engine provides
 * (AutoPasImpl.h:55-113; LogicHandler.h:58-100, checkMinimalSize :967-983). */
ljg_status ljg_create(const ljg_params *params, ljg_ctx **out);
ljg_status ljg_destroy(ljg_ctx *ctx);
const char *ljg_last_error(const ljg_ctx *ctx);
/* Human-readable name of the device the context runs on (for logs). */
ljg_status ljg_device_name(const ljg_ctx *ctx, char *buf, int32_t buflen);
/* Free and total device memory in bytes (hipMemGetInfo on the context's device): the HBM budget a
 * serving loop admits jobs against. */
ljg_status ljg_memory_info(const ljg_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* Change the periodic box of a context (AutoPas::setBoxMin/setBoxMax, AutoPasDecl.h:589-595, or
 * resizeBox, :116).  The particles are dropped: upload the next system afterwards.  Device buffers are
 * kept, so one context can serve many jobs of different sizes without reallocating. */
ljg_status ljg_set_box(ljg_ctx *ctx, const double box_min[3], const double box_max[3]);

/* Open (non-periodic) box with halo particles, the model of AutoPas's own containers: owned particles
 * lie in [box_min, box_max) and halo particles -- copies of other domains' particles or periodic images,
 * made by the caller as md-flexible's halo exchange does -- in the halo region up to rc + skin outside it
 * (AutoPas::addHaloParticle, AutoPasDecl.h:196; LinkedCells.h:87, VerletClusterLists.h:194).  Every pair
 * within rc interacts directly, without periodic images.  Particles of ljg_upload and
 * ljg_update_positions must lie in [box_min - (rc + skin), box_max + rc + skin]; others are refused
 * (LJG_ERR_ARG).  Forces come back for every particle.  ljg_step is refused: the caller advances the
 * particles and exchanges halos between force evaluations.  Every kernel runs on an open box (inside the
 * engine the box and its halo region sit in a periodic box with rc + skin of empty space beyond the halo,
 * so that no pair meets across the periodic boundary).  The particles are dropped, as by ljg_set_box,
 * which returns the context to a periodic box. */
ljg_status ljg_set_open_box(ljg_ctx *ctx, const double box_min[3], const double box_max[3]);

/* Ownership for the globals (AutoPas's OwnershipState, OwnershipState.h:20-30): the first n_owned particles
 * of the last upload are owned, the others halo.  Energy and virial then count as LJFunctor counts them
 * for owned i (LJFunctor.h:174-189): a pair of two owned particles whole, an owned-halo pair half, a
 * halo-halo pair not at all; the pair count still counts every pair.  ljg_upload resets it to every
 * particle owned.  With fewer owned than uploaded particles the Newton-3 kernels refuse globals
 * (LJG_ERR_UNSUPPORTED). */
ljg_status ljg_set_owned(ljg_ctx *ctx, int64_t n_owned);

/* Ensemble of independent periodic boxes in one context.  md-flexible runs one AutoPas instance per
 * simulation (AutoPasImpl.h:55-113); here the boxes share every launch.  Box b spans
 * [box_min[3b..3b+2], box_max[3b..3b+2]) and holds count[b] particles.  The next ljg_upload takes the
 * boxes' particles concatenated in box order (n = sum of count).  Every box side must be >= rc + skin.
 * Only LJG_KERNEL_CLUSTER_MIXED runs on an ensemble.  Each box keeps the tower grid, cluster layout, list
 * order and arithmetic it would have alone.  Lists are rebuilt for all boxes together: every
 * rebuild_frequency steps, or once a particle of any box moved skin/2.  ljg_set_box returns the context
 * to one box; the particles are dropped either way. */
ljg_status ljg_set_boxes(ljg_ctx *ctx, int32_t nboxes, const double *box_min, const double *box_max,
                         const int64_t *count);
/* Edit an ensemble in place, keeping running boxes resident on the device.  The new ensemble is the boxes
 * keep[0..nkeep) of the current one (indices into it, in their new order; positions, velocities,
 * forces, types and ids kept) followed by nadd new boxes given as for ljg_set_boxes + ljg_upload
 * (box_min/box_max 3 per box, count, and their particles concatenated; vx/vy/vz/type/id may be NULL).
 * Upload indices are renumbered box by box; the lists are rebuilt at the next force evaluation.  A
 * serving loop uses this to retire finished jobs and admit new ones between batched steps (md-flexible
 * starts and ends one simulation per AutoPas instance, AutoPasImpl.h:55-113). */
ljg_status ljg_edit_boxes(ljg_ctx *ctx, int32_t nkeep, const int32_t *keep, int32_t nadd, const double *box_min,
                          const double *box_max, const int64_t *count, const double *x, const double *y,
                          const double *z, const double *vx, const double *vy, const double *vz, const int32_t *type,
                          const int64_t *id);
/* Per-box results of the last ljg_force with calc_globals on an ensemble: potential energy and virial of
 * each box (LJFunctor::getPotentialEnergy / getVirial, LJFunctor.h:691-720) and its unique pairs within
 * rc (half its pair hits, rounded down, as ljg_force reports them for a box alone).  Any output pointer
 * may be NULL. */
ljg_status ljg_get_box_globals(ljg_ctx *ctx, int32_t nboxes, double *potential_energy, double *virial,
                               int64_t *pair_count);

/* Particle types: ParticlePropertiesLibrary::addSiteType + calculateMixingCoefficients
 * (ParticlePropertiesLibrary.h:443-489): Lorentz-Berthelot mixing is done here, in FP64. */
ljg_status ljg_set_types(ljg_ctx *ctx, int32_t ntypes, const double *epsilon, const double *sigma,
                         const double *mass);

/* Upload owned particles (AutoPas::addParticle, AutoPasDecl.h:172).  vx/vy/vz/type/id may be NULL
 * (zero / type 0 / index).  Positions outside the periodic box are wrapped like md-flexible's
 * RegularGridDecomposition.cpp:563-587.  Forces start at zero. */
ljg_status ljg_upload(ljg_ctx *ctx, int64_t n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz, const int32_t *type, const int64_t *id);
/* Set the current force (e.g. F0 restored from a checkpoint, MDFlexConfig.cpp:173). Upload order. */
ljg_status ljg_set_forces(ljg_ctx *ctx, const double *fx, const double *fy, const double *fz);
/* Overwrite the positions (and the velocities, if vx/vy/vz are given) of the uploaded particles, in
 * upload order, keeping the current layout and neighbour structure: the caller moved its particles between
 * two force evaluations, as md-flexible's TimeDiscretization does through AutoPas's iterators
 * (TimeDiscretization.cpp:17-69).  *moved (if given) is set to 1 when a particle is now skin/2 or more
 * from where it was at the last rebuild (LogicHandler.h:1000-1015), else 0. */
ljg_status ljg_update_positions(ljg_ctx *ctx, const double *x, const double *y, const double *z, const double *vx,
                                const double *vy, const double *vz, int32_t *moved);
/* Rebuild the neighbour structure the given kernel runs on, now (LogicHandler::rebuildNeighborLists,
 * LogicHandler.h:1082-1090): wrap into the box, bin or cluster, build lists. */
ljg_status ljg_rebuild(ljg_ctx *ctx, int32_t kernel);
/* Cluster size and tile shape of the next build (VerletClusterLists clusterSize, ljg_params.cluster_size,
 * tile_groups, build_groups): a tuner switching between GPU configurations (AutoTuner::tuneConfiguration,
 * AutoTuner.cpp:85-175) changes them between force evaluations; a different cluster size rebuilds the
 * lists at the next evaluation. */
ljg_status ljg_set_layout(ljg_ctx *ctx, int32_t cluster_size, int32_t tile_groups, int32_t build_groups);
/* Truncated-shifted energy on or off (LJFunctor's applyShift template argument, LJFunctor.h:39-44); the
 * mixing table is recomputed from the last ljg_set_types parameters. */
ljg_status ljg_set_shift(ljg_ctx *ctx, int32_t apply_shift);

/* LinkedCells binning (counting sort into CellBlock3D cells; CellBlock3D.h:320-425, LinkedCells.h:147-203).
 * Leaves the particle state in cell order.  Used by the plain FP64 kernel. */
ljg_status ljg_bin(ljg_ctx *ctx);
/* Verlet-cluster build (towers, z-sort, clusters, cluster-pair lists): VerletClusterLists::
 * rebuildNeighborLists (VerletClusterLists.h:779-800, VerletClusterListsRebuilder.h:67-409).
 * Leaves the particle state in cluster order. */
ljg_status ljg_build_pairs(ljg_ctx *ctx);

/* One force evaluation on the current positions with the given kernel: container->computeInteractions
 * (traversal) + LJFunctor::endTraversal (LJFunctor.h:661-685).  Builds the structure the kernel needs
 * if it is missing.  out may be NULL. */
ljg_status ljg_force(ljg_ctx *ctx, int32_t kernel, int32_t calc_globals, ljg_globals *out);

/* nsteps velocity-Verlet steps (TimeDiscretization.cpp:17-69, :152-165) with the rebuild policy of
 * LogicHandler.h:987-1015, all on the device.  out may be NULL. */
ljg_status ljg_step(ljg_ctx *ctx, int32_t nsteps, double dt, int32_t kernel, ljg_run_stats *out);

/* Layout tuning state and scores (AutoTuner's evidence, AutoTuner.cpp:250-369). */
ljg_status ljg_get_tuning(ljg_ctx *ctx, ljg_tuning *out);

/* Turn HIP-event timing of force kernels and rebuilds on/off (events on the context's stream). */
ljg_status ljg_set_timing(ljg_ctx *ctx, int32_t enabled);

/* Download state in upload order.  Any pointer may be NULL. */
ljg_status ljg_download(ljg_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                        double *fx, double *fy, double *fz);
/* Download particles [first, first + count) of the upload order (one box of an ensemble, for example). */
ljg_status ljg_download_range(ljg_ctx *ctx, int64_t first, int64_t count, double *x, double *y, double *z,
                              double *vx, double *vy, double *vz, double *fx, double *fy, double *fz);

/* ---- inspection (parity tests) -------------------------------------------------------------- */
/* Cell of each particle (upload order) after ljg_bin: CellBlock3D::get1DIndexOfPosition of the
 * position the particle had when it was binned. */
ljg_status ljg_get_cells(ljg_ctx *ctx, int64_t *cell_of_particle, int64_t cells_per_dim_with_halo[3]);
/* Sizes of the cluster structure after ljg_build_pairs. */
ljg_status ljg_cluster_sizes(ljg_ctx *ctx, int64_t *num_slots, int64_t *num_clusters, int64_t *num_entries,
                             int64_t towers_per_dim[2], double side[2], int64_t *towers_per_il);
/* slot_particle[num_slots] (upload index, -1 for padding), cluster_tower[num_clusters],
 * list_offsets[num_clusters+1], list_j[num_entries], list_shift[num_entries], bbox[6*num_clusters]
 * (xlo,ylo,zlo,xhi,yhi,zhi).  Any pointer may be NULL.  After a build for
 * LJG_KERNEL_CLUSTER_JROWS_MIXED the entries are j particles: list_j holds their upload index and
 * list_shift is -1 (minimum image). */
ljg_status ljg_get_clusters(ljg_ctx *ctx, int64_t *slot_particle, int32_t *cluster_tower, int64_t *list_offsets,
                            int32_t *list_j, int32_t *list_shift, double *bbox);

/* After a cluster-entry build with member rows (every one for LJG_KERNEL_CLUSTER_MIXED unless
 * whole_cluster_rows is set): per i-cluster the j-cluster members the FP32 cluster kernel streams, as
 * entry << 3 | member (entry: position in the cluster's entries of ljg_get_clusters).  offsets has
 * num_clusters + 1 entries; rows may be NULL to query the sizes (offsets[num_clusters] is the total).
 * Inspection entry point of this engine's layout (no AutoPas counterpart). */
ljg_status ljg_get_member_rows(ljg_ctx *ctx, int64_t *offsets, int32_t *rows);

/* Neighbour lists after a Verlet-list build, in upload indices: neighbours of particle p are
 * neighbors[offsets[p] .. offsets[p+1]).  offsets has n+1 entries; neighbors may be NULL to query sizes
 * (offsets[n] is the total). */
ljg_status ljg_get_verlet_lists(ljg_ctx *ctx, int64_t *offsets, int32_t *neighbors);

/* ---- allocation-free force evaluation ------------------------------------------------------------ */
/* Limits of ljg_force_lc_noalloc: its intermediates live in one static device array of
 * LJG_NOALLOC_SCRATCH_BYTES per GPU (held by the library until the process ends). */
#define LJG_NOALLOC_MAX_PARTICLES 524288
#define LJG_NOALLOC_MAX_CELLS 524288 /* CellBlock3D cells including the halo layers */
#define LJG_NOALLOC_SCRATCH_BYTES (48u << 20)
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the plain FP64 LinkedCells force
 * evaluation of ljg_upload + ljg_force(LJG_KERNEL_LC_FP64, calc_globals) without a context and without any
 * allocation.  x, y, z, type, id (inputs), fx, fy, fz and globals (outputs) are DEVICE pointers on the
 * current HIP device; type (values in [0, ntypes), clamped) and id may be NULL, as in ljg_upload, and
 * globals may be NULL (no energy, virial or pair count).  The box, cutoff, skin and apply_shift come from
 * params (the other fields are ignored); ntypes types with per-type epsilon and sigma are mixed as
 * ljg_set_types does (ntypes 1 with NULL arrays: epsilon = sigma = 1).  Positions are wrapped into the box as
 * ljg_upload does, without changing the caller's arrays.  Everything is queued on `stream` (a hipStream_t,
 * NULL for the default stream) and the call returns without waiting; fx, fy, fz and *globals are
 * bit-identical to the context's results.  Returns LJG_ERR_CAPACITY when n > LJG_NOALLOC_MAX_PARTICLES or
 * the box has more than LJG_NOALLOC_MAX_CELLS cells.  The scratch array is shared by every call on a GPU
 * and keeps the last call's contents (the call clears what must start at zero): calls that overlap on one
 * GPU (other streams or threads) race and must be serialised by the caller. */
ljg_status ljg_force_lc_noalloc(const ljg_params *params, int32_t ntypes, const double *epsilon, const double *sigma,
                                int64_t n, const double *x, const double *y, const double *z, const int32_t *type,
                                const int64_t *id, double *fx, double *fy, double *fz, ljg_globals *globals,
                                void *stream);

/* ---- checkpoints (host side; md-flexible's .pvtu/.vtu format) --------------------------------- */
/* Write dir/<name>_Particles_<iteration>.pvtu and dir/data/<name>_Particles_0_<iteration>.vtu, the
 * iteration zero-padded to `digits` (ParallelVtkWriter.cpp:60-200,309-358,438-442), with 17 significant
 * digits so FP64 state round-trips.  Arrays in caller order.  The index path is copied to pvtu_path if
 * given. */
ljg_status ljg_write_checkpoint(const char *dir, const char *name, int64_t iteration, int32_t digits, int64_t n,
                                const double *x, const double *y, const double *z, const double *vx,
                                const double *vy, const double *vz, const double *fx, const double *fy,
                                const double *fz, const int32_t *type, const int64_t *id, char *pvtu_path,
                                int32_t pvtu_path_len);
/* Read every piece of a checkpoint (MDFlexConfig::loadParticlesFromCheckpoint, MDFlexConfig.cpp:33-183,
 * 648-670).  With x == NULL only *n is set (the total); otherwise *n is the capacity of the arrays on
 * input and the count on output.  Pieces are concatenated in rank order. */
ljg_status ljg_read_checkpoint(const char *pvtu, int64_t *n, double *x, double *y, double *z, double *vx, double *vy,
                               double *vz, double *fx, double *fy, double *fz, int32_t *type, int64_t *id);

/* ---- synthetic systems (host-side, seeded Philox4x32-10) ------------------------------------- */
/* Simple-cubic lattice of nx*ny*nz sites, spacing a, positions box_min + (i+1/2)a plus uniform jitter in
 * [-jitter, jitter]; the first n sites are used (n <= nx*ny*nz).  Velocities Maxwell-Boltzmann at
 * temperature T for mass 1, net momentum removed, rescaled to T exactly with 3n-3 degrees of freedom.
 * Types are assigned uniformly in [0, ntypes) if ntypes > 1 (type may be NULL otherwise). */
ljg_status ljg_generate_lattice(uint64_t seed, int64_t n, int64_t nx, int64_t ny, int64_t nz, double a,
                                const double box_min[3], double jitter, double temperature, int32_t ntypes,
                                double *x, double *y, double *z, double *vx, double *vy, double *vz, int32_t *type);

/* Two-phase system: n_liquid particles uniform in the slab slab_lo <= z < slab_hi, then n_gas uniform
 * in the rest of the periodic box [0, box), placed one by one with no pair closer than dmin (random
 * sequential placement, periodic minimum image).  Types uniform in [0, ntypes); velocities as for the
 * lattice.  Returns LJG_ERR_CAPACITY if max_attempts candidates were not enough. */
ljg_status ljg_generate_slab(uint64_t seed, const double box[3], double slab_lo, double slab_hi, int64_t n_liquid,
                             int64_t n_gas, double dmin, int32_t ntypes, double temperature, int64_t max_attempts,
                             double *x, double *y, double *z, double *vx, double *vy, double *vz, int32_t *type);

#ifdef __cplusplus
}
#endif
#endif
