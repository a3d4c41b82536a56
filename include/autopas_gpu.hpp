/*
 * autopas_gpu.hpp -- C++ host API of the MI355X engine: the part of AutoPas's interface that the
 * Lennard-Jones short-range path uses, over the C ABI of ljgpu.h.  Header-only; link libljgpu.so.
 *
 * A program written against AutoPas for this path,
 *
 *   autopas::AutoPas<mdLib::MoleculeLJ> autoPas;            AutoPasDecl.h:46-110
 *   autoPas.setBoxMin(..); autoPas.setBoxMax(..); autoPas.setCutoff(..); autoPas.setVerletSkin(..);
 *   autoPas.setVerletRebuildFrequency(..); autoPas.init();
 *   autoPas.addParticle(p);                                 AutoPasDecl.h:172
 *   mdLib::LJFunctor<mdLib::MoleculeLJ, true, true, autopas::FunctorN3Modes::Both, true> f(cutoff, ppl);
 *   autoPas.computeInteractions(&f);                        AutoPasDecl.h:271, AutoPasImpl.h:116-135
 *   for (auto &p : autoPas) { ... }                         AutoPasDecl.h:280
 *
 * compiles against the namespace autopas_amd with the same calls.  What differs:
 *   - by default the box is periodic inside the engine: updateContainer() returns no leaving particles,
 *     and the positions come back wrapped into the box after each list rebuild (md-flexible's periodic
 *     wrap, RegularGridDecomposition.cpp:563-587, done on the device).  setPeriodicBoundaries(false)
 *     gives AutoPas's own model instead: an open box, halo particles added by the caller
 *     (addHaloParticle, e.g. md-flexible's halo exchange), and updateContainer() returning the owned
 *     particles that left the box and deleting the halo particles (LogicHandler.h:165-218);
 *   - begin()/end() iterate the owned particles (forEach and getNumberOfParticles take an
 *     IteratorBehavior);
 *   - the functor's per-pair and per-SoA entry points are not provided (the GPU evaluates the whole
 *     container); they throw "not implemented" like PairwiseFunctor's defaults (PairwiseFunctor.h:56-104);
 *   - step(n, dt) runs n velocity-Verlet steps on the device (the engine's own loop);
 *   - setKernel() picks the GPU layout; setKernel(LJG_KERNEL_AUTO) lets the AutoTuner below pick it among the
 *     allowed GPU configurations, as AutoPas picks containers and traversals.
 * Every error is thrown as autopas_amd::AutoPasException, as ExceptionHandler does by default
 * (ExceptionHandler.cpp:12).
 */
#ifndef AUTOPAS_GPU_HPP
#define AUTOPAS_GPU_HPP

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "ljgpu.h"

namespace autopas_amd {

/* autopas::utils::ExceptionHandler::AutoPasException (ExceptionHandler.h:116-140). */
class AutoPasException : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

namespace detail {
inline void check(const ljg_ctx *ctx, ljg_status st, const char *what) {
  if (st == LJG_OK) return;
  std::string msg = std::string(what) + " failed (status " + std::to_string(st) + ")";
  if (st == LJG_ERR_NODEVICE) msg += ": no HIP device available";
  if (ctx) {
    const char *e = ljg_last_error(ctx);
    if (e && *e) msg += ": " + std::string(e);
  }
  throw AutoPasException(msg);
}
}  // namespace detail

/* autopas::OwnershipState (OwnershipState.h:20-29). */
enum class OwnershipState : int64_t { dummy = 0, owned = 1, halo = 2 };

/* autopas::IteratorBehavior (IteratorBehavior.h:34-42): which particles an iteration or count covers. */
enum class IteratorBehavior : int { owned = 0b01, halo = 0b10, ownedOrHalo = 0b11 };

/* autopas::FunctorN3Modes (Functor.h). */
enum class FunctorN3Modes { Newton3Only, Newton3Off, Both };

/* mdLib::MoleculeLJ (MoleculeLJ.h:20-212 over ParticleBase.h:32-445): position, velocity, force, old
 * force, id, type and ownership. */
class MoleculeLJ {
 public:
  MoleculeLJ() = default;
  MoleculeLJ(const std::array<double, 3> &pos, const std::array<double, 3> &v, unsigned long moleculeId,
             unsigned long typeId = 0)
      : _r(pos), _v(v), _id(moleculeId), _typeId(typeId) {}

  const std::array<double, 3> &getR() const { return _r; }
  void setR(const std::array<double, 3> &r) { _r = r; }
  void addR(const std::array<double, 3> &r) {
    for (int d = 0; d < 3; ++d) _r[d] += r[d];
  }
  const std::array<double, 3> &getV() const { return _v; }
  void setV(const std::array<double, 3> &v) { _v = v; }
  void addV(const std::array<double, 3> &v) {
    for (int d = 0; d < 3; ++d) _v[d] += v[d];
  }
  const std::array<double, 3> &getF() const { return _f; }
  void setF(const std::array<double, 3> &f) { _f = f; }
  void addF(const std::array<double, 3> &f) {
    for (int d = 0; d < 3; ++d) _f[d] += f[d];
  }
  void subF(const std::array<double, 3> &f) {
    for (int d = 0; d < 3; ++d) _f[d] -= f[d];
  }
  const std::array<double, 3> &getOldF() const { return _oldF; }
  void setOldF(const std::array<double, 3> &f) { _oldF = f; }
  size_t getID() const { return _id; }
  void setID(size_t id) { _id = id; }
  size_t getTypeId() const { return _typeId; }
  void setTypeId(size_t t) { _typeId = t; }
  OwnershipState getOwnershipState() const { return _ownership; }
  void setOwnershipState(OwnershipState s) { _ownership = s; }
  bool isOwned() const { return _ownership == OwnershipState::owned; }
  bool isHalo() const { return _ownership == OwnershipState::halo; }
  bool isDummy() const { return _ownership == OwnershipState::dummy; }

 private:
  std::array<double, 3> _r{}, _v{}, _f{}, _oldF{};
  size_t _id = 0;
  size_t _typeId = 0;
  OwnershipState _ownership = OwnershipState::owned;
};

/* ParticlePropertiesLibrary<double, size_t> (ParticlePropertiesLibrary.h:30-489): per-site mass, epsilon
 * and sigma, Lorentz-Berthelot mixing (:443-489) and the shift (calcShift6, :575-582).  The engine mixes
 * the same way on its own from the per-site values; the tables here serve the getters. */
class ParticlePropertiesLibrary {
 public:
  explicit ParticlePropertiesLibrary(double cutoff) : _cutoff(cutoff) {}

  void addSiteType(size_t siteId, double mass) {
    if (siteId != _mass.size())
      throw AutoPasException("ParticlePropertiesLibrary::addSiteType(): trying to register a site type with id " +
                             std::to_string(siteId) +
                             ". Please register types consecutively, starting at id 0. Currently there are " +
                             std::to_string(_mass.size()) + " registered types.");
    _mass.push_back(mass);
    _eps.push_back(0.);
    _sig.push_back(0.);
  }
  void addLJParametersToSite(size_t siteId, double epsilon, double sigma) {
    if (siteId >= _mass.size())
      throw AutoPasException("ParticlePropertiesLibrary::addLJParametersToSite(): site id " +
                             std::to_string(siteId) + " is not registered.");
    _eps[siteId] = epsilon;
    _sig[siteId] = sigma;
  }
  void calculateMixingCoefficients() {
    const size_t n = _mass.size();
    if (n == 0) throw AutoPasException("ParticlePropertiesLibrary::calculateMixingCoefficients(): no site types.");
    const double rc2 = _cutoff * _cutoff;
    _mix.assign(3 * n * n, 0.);
    for (size_t a = 0; a < n; ++a) {
      for (size_t b = 0; b < n; ++b) {
        const double e24 = 24 * std::sqrt(_eps[a] * _eps[b]);
        const double s = (_sig[a] + _sig[b]) / 2.0;
        const double s2 = s * s;
        const double q2 = s2 / rc2, q6 = q2 * q2 * q2;
        double *m = &_mix[3 * (n * a + b)];
        m[0] = e24;
        m[1] = s2;
        m[2] = e24 * (q6 - q6 * q6);
      }
    }
  }
  int getNumberRegisteredSiteTypes() const { return (int)_mass.size(); }
  double getSiteMass(size_t i) const { return _mass.at(i); }
  double getEpsilon(size_t i) const { return _eps.at(i); }
  double getSigma(size_t i) const { return _sig.at(i); }
  double getMixing24Epsilon(size_t i, size_t j) const { return mixed(i, j, 0); }
  double getMixingSigmaSquared(size_t i, size_t j) const { return mixed(i, j, 1); }
  double getMixingShift6(size_t i, size_t j) const { return mixed(i, j, 2); }
  double getCutoff() const { return _cutoff; }

  const std::vector<double> &masses() const { return _mass; }
  const std::vector<double> &epsilons() const { return _eps; }
  const std::vector<double> &sigmas() const { return _sig; }

 private:
  double mixed(size_t i, size_t j, int k) const {
    const size_t n = _mass.size();
    if (_mix.size() != 3 * n * n)
      throw AutoPasException("ParticlePropertiesLibrary: calculateMixingCoefficients() was not called.");
    return _mix.at(3 * (n * i + j) + k);
  }
  double _cutoff;
  std::vector<double> _mass, _eps, _sig, _mix;
};

/* mdLib::LJFunctor<Particle_T, applyShift, useMixing, useNewton3, calculateGlobals, countFLOPs,
 * relevantForTuning> (LJFunctor.h:39-120).  The object passed to AutoPas::computeInteractions; the engine
 * evaluates the whole container with it and fills in what endTraversal would (LJFunctor.h:641-720). */
template <class Particle_T, bool applyShift = false, bool useMixing = false,
          FunctorN3Modes useNewton3 = FunctorN3Modes::Both, bool calculateGlobals = false, bool countFLOPs = false,
          bool relevantForTuning = true>
class LJFunctor {
 public:
  static constexpr bool kApplyShift = applyShift;
  static constexpr bool kUseMixing = useMixing;
  static constexpr bool kCalculateGlobals = calculateGlobals;
  static constexpr bool kCountFLOPs = countFLOPs;

  LJFunctor() = delete;
  explicit LJFunctor(double cutoff) : _cutoff(cutoff) {
    static_assert(not useMixing,
                  "Mixing without a ParticlePropertiesLibrary is not possible! Use a different constructor or set "
                  "mixing to false.");
  }
  LJFunctor(double cutoff, ParticlePropertiesLibrary &particlePropertiesLibrary)
      : _cutoff(cutoff), _ppl(&particlePropertiesLibrary) {
    static_assert(useMixing,
                  "Not using Mixing but using a ParticlePropertiesLibrary is not allowed! Use a different constructor "
                  "or set mixing to true.");
  }

  std::string getName() { return "LJFunctorGPU"; }
  bool isRelevantForTuning() { return relevantForTuning; }
  bool allowsNewton3() { return useNewton3 == FunctorN3Modes::Newton3Only or useNewton3 == FunctorN3Modes::Both; }
  bool allowsNonNewton3() { return useNewton3 == FunctorN3Modes::Newton3Off or useNewton3 == FunctorN3Modes::Both; }
  double getCutoff() const { return _cutoff; }

  /* Single-type constants (LJFunctor.h:588-596): 24 epsilon and sigma^2. */
  void setParticleProperties(double epsilon24, double sigmaSquared) {
    _epsilon24 = epsilon24;
    _sigmaSquared = sigmaSquared;
  }
  double getEpsilon24() const { return _epsilon24; }
  double getSigmaSquared() const { return _sigmaSquared; }
  ParticlePropertiesLibrary *getParticlePropertiesLibrary() const { return _ppl; }

  /* LJFunctor.h:641-655: globals and FLOP counters start again. */
  void initTraversal() {
    _postProcessed = false;
    _potentialEnergy = 0.;
    _virial = 0.;
    _pairsNoN3 = _pairsN3 = _distanceCalls = 0;
  }

  /* LJFunctor.h:691-720, same messages. */
  double getPotentialEnergy() {
    if (not calculateGlobals)
      throw AutoPasException(
          "Trying to get potential energy even though calculateGlobals is false. If you want this functor to "
          "calculate global values, please specify calculateGlobals to be true.");
    if (not _postProcessed) throw AutoPasException("Cannot get potential energy, because endTraversal was not called.");
    return _potentialEnergy;
  }
  double getVirial() {
    if (not calculateGlobals)
      throw AutoPasException(
          "Trying to get virial even though calculateGlobals is false. If you want this functor to calculate "
          "global values, please specify calculateGlobals to be true.");
    if (not _postProcessed) throw AutoPasException("Cannot get virial, because endTraversal was not called.");
    return _virial;
  }

  /* FLOP model of LJFunctor.h:722-789: 8 per distance, 15 per non-Newton-3 kernel call and 18 per Newton-3
   * one, 8 or 9 per non-Newton-3 globals calculation and 12 or 13 per Newton-3 one.  A full-list kernel
   * calls the pair kernel twice per pair within rc (once from each side), a Newton-3 kernel once. */
  size_t getNumFLOPs() const {
    if constexpr (countFLOPs) {
      const size_t noN3 = 2 * _pairsNoN3, n3 = _pairsN3;
      size_t flops = 8 * _distanceCalls + 15 * noN3 + 18 * n3;
      if (calculateGlobals) flops += (applyShift ? 9 : 8) * noN3 + (applyShift ? 13 : 12) * n3;
      return flops;
    } else {
      return std::numeric_limits<size_t>::max();
    }
  }
  /* Kernel calls (pairs within rc) per distance calculation (LJFunctor.h:794-816). */
  double getHitRate() const {
    if constexpr (countFLOPs) {
      return _distanceCalls ? (double)(2 * _pairsNoN3 + _pairsN3) / (double)_distanceCalls : 0.;
    } else {
      return std::numeric_limits<double>::quiet_NaN();
    }
  }

  void AoSFunctor(Particle_T &, Particle_T &, bool) { throw AutoPasException(getName() + "::AoSFunctor: not implemented"); }
  template <class SoAView>
  void SoAFunctorSingle(SoAView, bool) {
    throw AutoPasException(getName() + "::SoAFunctorSingle: not implemented");
  }
  template <class SoAView>
  void SoAFunctorPair(SoAView, SoAView, bool) {
    throw AutoPasException(getName() + "::SoAFunctorPair: not implemented");
  }
  template <class SoAView, class NeighborList>
  void SoAFunctorVerlet(SoAView, size_t, const NeighborList &, bool) {
    throw AutoPasException(getName() + "::SoAFunctorVerlet: not implemented");
  }

  /* Called by AutoPas::computeInteractions with the engine's results (endTraversal's role); newton3: the
   * kernel evaluated each pair once. */
  void setTraversalResults(const ljg_globals &g, bool newton3 = false) {
    _potentialEnergy = g.potential_energy;
    _virial = g.virial;
    (newton3 ? _pairsN3 : _pairsNoN3) += (size_t)g.pair_count;
    _distanceCalls += g.candidate_evals > 0 ? (size_t)g.candidate_evals : 0;
    _postProcessed = true;
  }

 private:
  double _cutoff;
  ParticlePropertiesLibrary *_ppl = nullptr;
  double _epsilon24 = 24.;
  double _sigmaSquared = 1.;
  bool _postProcessed = false;
  double _potentialEnergy = 0., _virial = 0.;
  size_t _pairsNoN3 = 0, _pairsN3 = 0, _distanceCalls = 0;
};

/* ---- the GPU configurations in AutoPas's tuning space -------------------------------------------------
 * AutoPas tunes over Configurations (container, traversal, load estimator, data layout, newton3, cell size
 * factor; Configuration.h:26-171), the cartesian product of the allowed options filtered by compatibility
 * (SearchSpaceGenerators.cpp:18-67).  The engine enters that space as new container and traversal option
 * values, the Newton-3 option, and two GPU parameters: the precision of the pair term (the sums are FP64
 * in every kernel) and the cluster size with the i-clusters per wavefront of the force kernel. */
namespace options {
/* ContainerOption / TraversalOption entries for the engine (ContainerOption.h, TraversalOption.h:25-199). */
enum class ContainerOption { gpuLinkedCells, gpuVerletLists, gpuClusterLists };
enum class TraversalOption { gpu_lc_full_shell, gpu_verlet_lists, gpu_cluster_tiles, gpu_cluster_jrows };
/* autopas::Newton3Option (Newton3Option.h). */
enum class Newton3Option { disabled, enabled };
/* GPU parameter: precision of the pair term. */
enum class PrecisionOption { fp64, fp32pair };

inline const char *toString(ContainerOption o) {
  switch (o) {
    case ContainerOption::gpuLinkedCells: return "gpuLinkedCells";
    case ContainerOption::gpuVerletLists: return "gpuVerletLists";
    default: return "gpuClusterLists";
  }
}
inline const char *toString(TraversalOption o) {
  switch (o) {
    case TraversalOption::gpu_lc_full_shell: return "gpu_lc_full_shell";
    case TraversalOption::gpu_verlet_lists: return "gpu_verlet_lists";
    case TraversalOption::gpu_cluster_tiles: return "gpu_cluster_tiles";
    default: return "gpu_cluster_jrows";
  }
}
}  // namespace options

/* autopas::Configuration (Configuration.h:26-171) with the engine's parameters.  clusterSize and
 * tileGroups are 0 where they do not apply (as the load estimator is "none" where it does not). */
struct Configuration {
  options::ContainerOption container = options::ContainerOption::gpuClusterLists;
  options::TraversalOption traversal = options::TraversalOption::gpu_cluster_jrows;
  options::Newton3Option newton3 = options::Newton3Option::disabled;
  options::PrecisionOption precision = options::PrecisionOption::fp32pair;
  int clusterSize = 0;
  int tileGroups = 0;  // i-clusters per wavefront of the force kernel; 0: the engine's default

  auto key() const { return std::make_tuple(container, traversal, newton3, precision, clusterSize, tileGroups); }
  bool operator<(const Configuration &o) const { return key() < o.key(); }
  bool operator==(const Configuration &o) const { return key() == o.key(); }
  bool operator!=(const Configuration &o) const { return !(*this == o); }

  /* The engine kernel of this configuration (LJG_KERNEL_*), -1 when there is none. */
  int32_t kernel() const {
    using namespace options;
    const bool n3 = newton3 == Newton3Option::enabled, f64 = precision == PrecisionOption::fp64;
    if (container == ContainerOption::gpuLinkedCells && traversal == TraversalOption::gpu_lc_full_shell)
      return f64 && !n3 ? LJG_KERNEL_LC_FP64 : -1;
    if (container == ContainerOption::gpuVerletLists && traversal == TraversalOption::gpu_verlet_lists)
      return !f64 && !n3 ? LJG_KERNEL_VERLET_MIXED : -1;
    if (container != ContainerOption::gpuClusterLists) return -1;
    if (traversal == TraversalOption::gpu_cluster_tiles)
      return f64 ? (n3 ? LJG_KERNEL_CLUSTER_FP64_N3 : LJG_KERNEL_CLUSTER_FP64)
                 : (n3 ? LJG_KERNEL_CLUSTER_MIXED_N3 : LJG_KERNEL_CLUSTER_MIXED);
    if (traversal == TraversalOption::gpu_cluster_jrows) return !f64 && !n3 ? LJG_KERNEL_CLUSTER_JROWS_MIXED : -1;
    return -1;
  }

  /* Configuration::hasCompatibleValues (Configuration.h): a kernel exists, the cluster parameters are set
   * exactly where they apply, and the tile shape is one the kernel is built for (0 = its default). */
  bool hasCompatibleValues() const {
    const int32_t k = kernel();
    if (k < 0) return false;
    const bool cl = container == options::ContainerOption::gpuClusterLists;
    if (!cl) return clusterSize == 0 && tileGroups == 0;
    if (clusterSize != 1 && clusterSize != 2 && clusterSize != 4 && clusterSize != 8) return false;
    if (tileGroups == 0) return true;
    if (k == LJG_KERNEL_CLUSTER_FP64 || k == LJG_KERNEL_CLUSTER_FP64_N3) return false;  // one i-cluster per wave
    static const std::map<std::pair<int, int>, std::set<int>> tiles = {
        {{LJG_KERNEL_CLUSTER_MIXED, 8}, {1, 2, 4, 8}},     {{LJG_KERNEL_CLUSTER_MIXED, 4}, {1, 2, 4, 8}},
        {{LJG_KERNEL_CLUSTER_MIXED, 2}, {1, 2, 4, 8, 16}}, {{LJG_KERNEL_CLUSTER_MIXED, 1}, {4, 8, 16, 32, 64}},
        {{LJG_KERNEL_CLUSTER_JROWS_MIXED, 8}, {2, 4, 8}},  {{LJG_KERNEL_CLUSTER_JROWS_MIXED, 4}, {4, 8, 16}},
        {{LJG_KERNEL_CLUSTER_JROWS_MIXED, 2}, {8, 16}},    {{LJG_KERNEL_CLUSTER_JROWS_MIXED, 1}, {16, 32}}};
    const int kk = k == LJG_KERNEL_CLUSTER_MIXED_N3 ? LJG_KERNEL_CLUSTER_MIXED : k;
    const auto it = tiles.find({kk, clusterSize});
    return it != tiles.end() && it->second.count(tileGroups) > 0;
  }

  /* Configuration::toShortString. */
  std::string toString() const {
    return std::string("{") + options::toString(container) + " , " + options::toString(traversal) + " , " +
           (newton3 == options::Newton3Option::enabled ? "N3" : "noN3") + " , " +
           (precision == options::PrecisionOption::fp64 ? "fp64" : "fp32pair") + " , M" + std::to_string(clusterSize) +
           " , G" + std::to_string(tileGroups) + "}";
  }
};

/* SearchSpaceGenerators::cartesianProduct (SearchSpaceGenerators.cpp:18-67): every combination of the
 * allowed values that has compatible values; the cluster parameters collapse to 0 where they do not
 * apply.  Throws when nothing is left, as the reference does. */
inline std::set<Configuration> cartesianProduct(const std::set<options::ContainerOption> &containers,
                                                const std::set<options::TraversalOption> &traversals,
                                                const std::set<options::Newton3Option> &newton3,
                                                const std::set<options::PrecisionOption> &precisions,
                                                const std::set<int> &clusterSizes, const std::set<int> &tileGroups) {
  std::set<Configuration> space;
  for (auto c : containers)
    for (auto t : traversals)
      for (auto n : newton3)
        for (auto p : precisions)
          for (int m : clusterSizes)
            for (int g : tileGroups) {
              Configuration cfg{c, t, n, p, m, g};
              if (c != options::ContainerOption::gpuClusterLists) cfg.clusterSize = cfg.tileGroups = 0;
              if (cfg.hasCompatibleValues()) space.insert(cfg);
            }
  if (space.empty()) throw AutoPasException("No valid configurations could be created.");
  return space;
}

/* autopas::AutoTuner (AutoTuner.h:40-443) with the full-search strategy and the fastestAbs selector
 * (AutoTunerInfo.h: maxSamples 3, tuningInterval 5000): a tuning phase starts every tuningInterval
 * iterations; it takes the configurations from the back of the queue, maxSamples iterations each, and
 * scores each by estimateRuntimeFromSamples (AutoTuner.cpp:360-369): the fastest rebuilding sample /
 * rebuild frequency + the fastest traversal sample.  The lowest score is used until the next phase.
 * Configurations the engine cannot run on this box are dropped (LogicHandler's reject loop,
 * LogicHandler.h:1212-1230). */
class AutoTuner {
 public:
  AutoTuner(std::set<Configuration> space, unsigned int maxSamples, unsigned int tuningInterval,
            unsigned int rebuildFrequency)
      : _space(std::move(space)),
        _maxSamples(maxSamples == 0 ? 1 : maxSamples),
        _tuningInterval(tuningInterval),
        _rebuildFrequency(rebuildFrequency == 0 ? 1 : rebuildFrequency) {
    if (_space.empty()) throw AutoPasException("AutoTuner: empty search space");
    _current = *_space.rbegin();
  }

  /* AutoTuner::tuneConfiguration (AutoTuner.cpp:85-175) for the iteration about to run; returns whether a
   * tuning phase is in progress (the iteration's measurement is then expected). */
  bool tuneConfiguration(size_t iteration) {
    if (!_isTuning && (iteration == 0 || (_tuningInterval > 0 && iteration - _phaseEnd >= _tuningInterval))) {
      _isTuning = true;
      ++_tuningPhase;
      _queue.assign(_space.begin(), _space.end());  // the back is tested first (FILO)
      _evidence.clear();
      resetSamples();
    }
    if (!_isTuning) return false;
    if (_traverse.size() >= _maxSamples) {
      _evidence[_queue.back()] = estimate();
      _queue.pop_back();
      resetSamples();
    }
    if (_queue.empty()) {
      selectBest(iteration);
      return false;
    }
    _current = _queue.back();
    return true;
  }

  /* AutoTuner::addMeasurement (AutoTuner.cpp:250-330): rebuild samples only from iterations that rebuilt. */
  void addMeasurement(long rebuildNs, long traverseNs, bool rebuilt) {
    if (!_isTuning) return;
    if (rebuilt) _rebuild.push_back(rebuildNs);
    _traverse.push_back(traverseNs);
  }

  /* The current configuration cannot run here: drop it from the space and the queue. */
  void rejectConfiguration(const Configuration &cfg) {
    _space.erase(cfg);
    _queue.erase(std::remove(_queue.begin(), _queue.end(), cfg), _queue.end());
    resetSamples();
    if (_space.empty()) throw AutoPasException("AutoTuner: no configuration is applicable");
    if (!_isTuning && cfg == _current) _current = *_space.rbegin();
  }

  const Configuration &getCurrentConfig() const { return _current; }
  bool inTuningPhase() const { return _isTuning; }
  size_t getTuningPhase() const { return _tuningPhase; }
  const std::set<Configuration> &getSearchSpace() const { return _space; }
  /* Effective time (ns) of every configuration scored in the last tuning phase. */
  const std::map<Configuration, long> &getEvidence() const { return _evidence; }

 private:
  void resetSamples() {
    _rebuild.clear();
    _traverse.clear();
  }
  long estimate() const {
    const long reb = _rebuild.empty() ? 0 : *std::min_element(_rebuild.begin(), _rebuild.end());
    const long trav = *std::min_element(_traverse.begin(), _traverse.end());
    return static_cast<long>(reb / _rebuildFrequency) + trav;
  }
  void selectBest(size_t iteration) {
    _isTuning = false;
    _phaseEnd = iteration;
    auto best = std::min_element(_evidence.begin(), _evidence.end(),
                                 [](const auto &a, const auto &b) { return a.second < b.second; });
    if (best != _evidence.end()) _current = best->first;
  }

  std::set<Configuration> _space;
  unsigned int _maxSamples, _tuningInterval, _rebuildFrequency;
  std::vector<Configuration> _queue;
  std::map<Configuration, long> _evidence;
  std::vector<long> _rebuild, _traverse;
  Configuration _current;
  bool _isTuning = false;
  size_t _tuningPhase = 0, _phaseEnd = 0;
};

/* autopas::AutoPas<Particle_T> (AutoPasDecl.h) for the Lennard-Jones path: owned particles (and, in an open
 * box, halo particles) kept host-side in insertion order and mirrored on the device. */
template <class Particle_T>
class AutoPas {
 public:
  using ParticleType = Particle_T;
  using iterator = typename std::vector<Particle_T>::iterator;
  using const_iterator = typename std::vector<Particle_T>::const_iterator;

  AutoPas() = default;
  AutoPas(const AutoPas &) = delete;
  AutoPas &operator=(const AutoPas &) = delete;
  ~AutoPas() {
    if (_ctx) ljg_destroy(_ctx);
  }

  void setBoxMin(const std::array<double, 3> &boxMin) { _boxMin = boxMin; }
  void setBoxMax(const std::array<double, 3> &boxMax) { _boxMax = boxMax; }
  const std::array<double, 3> &getBoxMin() const { return _boxMin; }
  const std::array<double, 3> &getBoxMax() const { return _boxMax; }
  void setCutoff(double cutoff) {
    if (cutoff <= 0.) throw AutoPasException("Cutoff <= 0.0: " + std::to_string(cutoff));
    _cutoff = cutoff;
  }
  double getCutoff() const { return _cutoff; }
  void setVerletSkin(double skin) { _skin = skin; }
  double getVerletSkin() const { return _skin; }
  void setVerletRebuildFrequency(unsigned int f) { _rebuildFrequency = f; }
  unsigned int getVerletRebuildFrequency() const { return _rebuildFrequency; }
  void setVerletClusterSize(unsigned int m) { _clusterSize = m; }
  unsigned int getVerletClusterSize() const { return _clusterSize; }
  /* true (the default): the box is periodic inside the engine.  false: an open box with caller-made halo
   * particles, as AutoPas's containers hold them (takes effect at init() or resizeBox()). */
  void setPeriodicBoundaries(bool periodic) { _periodic = periodic; }
  bool getPeriodicBoundaries() const { return _periodic; }

  /* GPU layout (LJG_KERNEL_*); -1 (default): the j-row kernel when every box side exceeds 2 (rc + skin),
   * else the FP32 cluster kernel.  LJG_KERNEL_AUTO: computeInteractions picks the configuration with the
   * AutoTuner over the allowed options below, and step() with the engine's own tuning phases. */
  void setKernel(int32_t kernel) { _kernel = kernel; }

  /* The tuning space of LJG_KERNEL_AUTO (AutoPas::setAllowedContainers / setAllowedTraversals /
   * setAllowedNewton3Options, AutoPasDecl.h:800-1000, plus the GPU parameters), and the tuner's settings
   * (setTuningInterval, setMaxSamples). */
  void setAllowedContainers(const std::set<options::ContainerOption> &v) { _allowedContainers = v; }
  void setAllowedTraversals(const std::set<options::TraversalOption> &v) { _allowedTraversals = v; }
  void setAllowedNewton3Options(const std::set<options::Newton3Option> &v) { _allowedNewton3 = v; }
  void setAllowedPrecisions(const std::set<options::PrecisionOption> &v) { _allowedPrecisions = v; }
  void setAllowedClusterSizes(const std::set<int> &v) { _allowedClusterSizes = v; }
  void setAllowedTileGroups(const std::set<int> &v) { _allowedTileGroups = v; }
  void setTuningInterval(unsigned int v) { _tuningInterval = v; }
  void setMaxSamples(unsigned int v) { _maxSamples = v; }
  /* The configuration of the last computeInteractions (AutoPas::getCurrentConfigs). */
  const Configuration &getCurrentConfig() const { return _applied; }
  /* The tuner of LJG_KERNEL_AUTO (null before the first computeInteractions with it). */
  const AutoTuner *getAutoTuner() const { return _tuner.get(); }
  int32_t getKernel() const { return _kernel; }
  void setDevice(int32_t device) { _device = device; }

  /* AutoPas::init (AutoPasImpl.h:55-113): creates the engine context on the device. */
  void init() {
    if (_ctx) {
      ljg_destroy(_ctx);
      _ctx = nullptr;
    }
    ljg_params p{};
    for (int d = 0; d < 3; ++d) {
      p.box_min[d] = _boxMin[d];
      p.box_max[d] = _boxMax[d];
    }
    p.cutoff = _cutoff;
    p.skin = _skin;
    p.cluster_size = (int32_t)_clusterSize;
    p.rebuild_frequency = (int32_t)_rebuildFrequency;
    p.dynamic_rebuild = 1;
    p.apply_shift = 0;
    p.device = _device;
    ljg_ctx *ctx = nullptr;
    detail::check(nullptr, ljg_create(&p, &ctx), "AutoPas::init: ljg_create");
    _ctx = ctx;
    if (!_periodic)
      detail::check(_ctx, ljg_set_open_box(_ctx, _boxMin.data(), _boxMax.data()), "AutoPas::init: open box");
    _shift = 0;
    _structureChanged = true;
  }

  /* AutoPas::resizeBox (AutoPasDecl.h:116): particles outside the new box are removed and returned. */
  std::vector<Particle_T> resizeBox(const std::array<double, 3> &boxMin, const std::array<double, 3> &boxMax) {
    requireInit("resizeBox");
    detail::check(_ctx,
                  _periodic ? ljg_set_box(_ctx, boxMin.data(), boxMax.data())
                            : ljg_set_open_box(_ctx, boxMin.data(), boxMax.data()),
                  "AutoPas::resizeBox");
    _boxMin = boxMin;
    _boxMax = boxMax;
    std::vector<Particle_T> kept, left;
    for (auto &p : _particles) (inBox(p.getR()) ? kept : left).push_back(p);
    _particles = std::move(kept);
    _structureChanged = true;
    return left;
  }

  /* AutoPas::addParticle (AutoPasDecl.h:172, LogicHandler.h:339-360). */
  void addParticle(const Particle_T &p) {
    if (!inBox(p.getR()))
      throw AutoPasException("LogicHandler: Trying to add a particle that is not in the bounding box.");
    _particles.push_back(p);
    _particles.back().setOwnershipState(OwnershipState::owned);
    _structureChanged = true;
  }
  /* AutoPas::addHaloParticle (AutoPasDecl.h:203, LogicHandler.h:366-391): a copy of another domain's
   * particle or a periodic image, outside the box and within rc + skin of it.  Open boxes only: a periodic
   * box makes its images on the device. */
  void addHaloParticle(const Particle_T &p) {
    if (_periodic)
      throw AutoPasException("AutoPas::addHaloParticle: the box is periodic (setPeriodicBoundaries(false) for halo "
                             "particles)");
    if (inBox(p.getR()))
      throw AutoPasException("LogicHandler: Trying to add a halo particle that is not outside the box of the container.");
    _halo.push_back(p);
    _halo.back().setOwnershipState(OwnershipState::halo);
  }
  void deleteAllParticles() {
    _particles.clear();
    _halo.clear();
    _structureChanged = true;
  }
  /* AutoPas::getNumberOfParticles (AutoPasDecl.h:559). */
  size_t getNumberOfParticles(IteratorBehavior behavior = IteratorBehavior::owned) const {
    return (((int)behavior & (int)IteratorBehavior::owned) ? _particles.size() : 0) +
           (((int)behavior & (int)IteratorBehavior::halo) ? _halo.size() : 0);
  }

  /* AutoPas::updateContainer (AutoPasDecl.h:140, LogicHandler.h:165-218).  Periodic box: no particle
   * leaves, and positions are wrapped into the box on the device at the next list rebuild.  Open box: the
   * owned particles outside the box are removed and returned, and every halo particle is deleted. */
  std::vector<Particle_T> updateContainer() {
    if (_periodic) return {};
    std::vector<Particle_T> kept, left;
    for (auto &p : _particles) (inBox(p.getR()) ? kept : left).push_back(p);
    if (!left.empty()) _structureChanged = true;
    _particles = std::move(kept);
    _halo.clear();
    return left;
  }

  /* AutoPas::computeInteractions (AutoPasImpl.h:116-135, LogicHandler.h:1066-1143): the functor's pair
   * forces are added to the particles' forces, and the functor's globals are set.  Lists are rebuilt as
   * LogicHandler does (LogicHandler.h:987-1015): on the first call after a change of the particle set,
   * every getVerletRebuildFrequency() calls, and when a particle moved skin/2.  Returns false (no tuning
   * phase in progress). */
  template <class Functor>
  bool computeInteractions(Functor *f) {
    requireInit("computeInteractions");
    if (!f) throw AutoPasException("AutoPas::computeInteractions: null functor");
    if (f->getCutoff() != _cutoff)
      throw AutoPasException("AutoPas::computeInteractions: the functor's cutoff differs from AutoPas's cutoff");
    applyFunctor(*f);
    const size_t n = _particles.size();
    if (n == 0) {
      f->initTraversal();
      f->setTraversalResults(ljg_globals{});
      return false;
    }
    pushState();
    const bool autoTune = _kernel == LJG_KERNEL_AUTO;
    bool tuning = false;
    ljg_globals g{};
    int32_t kernelUsed = LJG_KERNEL_CLUSTER_MIXED;
    for (;;) {
      int32_t kernel = forceKernel();
      bool configChanged = false;
      if (autoTune) {
        // LogicHandler::selectConfiguration (LogicHandler.h:1169-1231): the tuner's configuration for this
        // iteration; a new one rebuilds the lists (a new container and traversal)
        if (!_tuner)
          _tuner.reset(new AutoTuner(cartesianProduct(_allowedContainers, _allowedTraversals, _allowedNewton3,
                                                      _allowedPrecisions, _allowedClusterSizes, _allowedTileGroups),
                                     _maxSamples, _tuningInterval, _rebuildFrequency));
        tuning = _tuner->tuneConfiguration(_iterations);
        const Configuration cfg = _tuner->getCurrentConfig();
        kernel = cfg.kernel();
        configChanged = !_haveApplied || cfg != _applied;
        if (configChanged) {
          const int m = cfg.clusterSize > 0 ? cfg.clusterSize : (int)_clusterSize;
          detail::check(_ctx, ljg_set_layout(_ctx, m, cfg.tileGroups, 0), "AutoPas::computeInteractions: layout");
          _applied = cfg;
          _haveApplied = true;
        }
      }
      const auto t0 = std::chrono::steady_clock::now();
      bool rebuilt = false;
      if (configChanged || _rebuildDue || _stepsSinceRebuild >= _rebuildFrequency) {
        const ljg_status st = ljg_rebuild(_ctx, kernel);
        if (autoTune && st == LJG_ERR_UNSUPPORTED) {  // not applicable to this box: the reject loop
          _tuner->rejectConfiguration(_applied);
          _haveApplied = false;
          continue;
        }
        detail::check(_ctx, st, "AutoPas::computeInteractions: rebuild");
        _stepsSinceRebuild = 0;
        _rebuildDue = false;
        ++_numRebuilds;
        rebuilt = true;
      }
      const auto t1 = std::chrono::steady_clock::now();
      f->initTraversal();
      detail::check(_ctx, ljg_force(_ctx, kernel, Functor::kCalculateGlobals ? 1 : 0, &g),
                    "AutoPas::computeInteractions: ljg_force");
      kernelUsed = kernel;
      const auto t2 = std::chrono::steady_clock::now();
      // LogicHandler.h:1309-1324: the iteration's rebuild and traversal times go to the tuner
      if (tuning)
        _tuner->addMeasurement(std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count(),
                               std::chrono::duration_cast<std::chrono::nanoseconds>(t2 - t1).count(), rebuilt);
      break;
    }
    f->setTraversalResults(g, kernelUsed == LJG_KERNEL_CLUSTER_FP64_N3 || kernelUsed == LJG_KERNEL_CLUSTER_MIXED_N3);
    pullState(true);
    ++_stepsSinceRebuild;
    ++_iterations;
    return tuning;
  }

  /* AutoPas::getMeanRebuildFrequency (AutoPasDecl.h:1111, LogicHandler.h:633-646): iterations per list
   * rebuild so far (computeInteractions calls and device steps), or the configured frequency before the
   * first rebuild. */
  double getMeanRebuildFrequency() const {
    if (_numRebuilds == 0) return static_cast<double>(_rebuildFrequency);
    return static_cast<double>(_iterations) / static_cast<double>(_numRebuilds);
  }
  /* List rebuilds so far (LogicHandler's _numRebuilds, LogicHandler.h:839,1102). */
  size_t getNumRebuilds() const { return _numRebuilds; }

  /* The engine's own loop: nsteps velocity-Verlet steps on the device (TimeDiscretization.cpp:17-165 with
   * the rebuild policy of LogicHandler.h:987-1015), then the particles are brought back. */
  ljg_run_stats step(int32_t nsteps, double dt) {
    requireInit("step");
    ljg_run_stats rs{};
    if (_particles.empty() || nsteps <= 0) return rs;
    pushState();
    const int32_t kernel = _kernel >= 0 ? _kernel : defaultKernel();
    detail::check(_ctx, ljg_step(_ctx, nsteps, dt, kernel, &rs), "AutoPas::step");
    pullState(false);
    _stepsSinceRebuild = 0;
    _rebuildDue = false;
    _numRebuilds += static_cast<size_t>(rs.rebuilds);
    _iterations += static_cast<size_t>(rs.steps);
    return rs;
  }

  iterator begin() { return _particles.begin(); }
  iterator end() { return _particles.end(); }
  const_iterator begin() const { return _particles.begin(); }
  const_iterator end() const { return _particles.end(); }
  /* AutoPas::forEach (AutoPasDecl.h:318): owned particles, then halo particles. */
  template <class Lambda>
  void forEach(Lambda &&fn, IteratorBehavior behavior = IteratorBehavior::ownedOrHalo) {
    if ((int)behavior & (int)IteratorBehavior::owned)
      for (auto &p : _particles) fn(p);
    if ((int)behavior & (int)IteratorBehavior::halo)
      for (auto &p : _halo) fn(p);
  }

  ljg_ctx *context() { return _ctx; }

 private:
  void requireInit(const char *what) const {
    if (!_ctx) throw AutoPasException(std::string("AutoPas::") + what + ": init() was not called");
  }
  bool inBox(const std::array<double, 3> &r) const {
    for (int d = 0; d < 3; ++d)
      if (!(r[d] >= _boxMin[d] && r[d] < _boxMax[d])) return false;
    return true;
  }
  int32_t defaultKernel() const {
    for (int d = 0; d < 3; ++d)
      if (!(_boxMax[d] - _boxMin[d] > 2. * (_cutoff + _skin))) return LJG_KERNEL_CLUSTER_MIXED;
    return LJG_KERNEL_CLUSTER_JROWS_MIXED;
  }
  int32_t forceKernel() const {
    if (_kernel == LJG_KERNEL_AUTO || _kernel < 0) return defaultKernel();
    return _kernel;
  }

  template <class Functor>
  void applyFunctor(Functor &f) {
    std::vector<double> eps, sig, mass;
    if constexpr (Functor::kUseMixing) {
      const ParticlePropertiesLibrary *ppl = f.getParticlePropertiesLibrary();
      eps = ppl->epsilons();
      sig = ppl->sigmas();
      mass = ppl->masses();
      for (const auto *v : {&_particles, &_halo})
        for (const auto &p : *v)
          if (p.getTypeId() >= eps.size())
            throw AutoPasException("AutoPas::computeInteractions: particle type " + std::to_string(p.getTypeId()) +
                                   " is not registered in the ParticlePropertiesLibrary");
    } else {
      eps = {f.getEpsilon24() / 24.};
      sig = {std::sqrt(f.getSigmaSquared())};
      mass = {1.};
      for (const auto *v : {&_particles, &_halo})
        for (const auto &p : *v)
          if (p.getTypeId() != 0)
            throw AutoPasException("AutoPas::computeInteractions: a functor without mixing needs type 0 particles");
    }
    if (eps != _eps || sig != _sig || mass != _mass) {
      detail::check(_ctx, ljg_set_types(_ctx, (int32_t)eps.size(), eps.data(), sig.data(), mass.data()),
                    "AutoPas::computeInteractions: ljg_set_types");
      _eps = eps;
      _sig = sig;
      _mass = mass;
    }
    const int32_t shift = Functor::kApplyShift ? 1 : 0;
    if (shift != _shift) {
      detail::check(_ctx, ljg_set_shift(_ctx, shift), "AutoPas::computeInteractions: ljg_set_shift");
      _shift = shift;
    }
  }

  /* Host particles -> device: a full upload after a change of the particle set, else positions and
   * velocities into the current layout (which also tells whether a particle moved skin/2).  An open box
   * uploads the owned particles followed by the halo particles; the halo set counts as unchanged while the
   * ids come in the same order (md-flexible re-adds the same halo copies every iteration between
   * rebuilds, as AutoPas's updateHaloParticle expects, LogicHandler.h:380-388). */
  void pushState() {
    const size_t no = _particles.size(), n = no + _halo.size();
    auto at = [&](size_t i) -> const Particle_T & { return i < no ? _particles[i] : _halo[i - no]; };
    std::vector<double> a(6 * n);
    std::vector<int64_t> id(n);
    for (size_t i = 0; i < n; ++i) {
      const auto &p = at(i);
      for (int d = 0; d < 3; ++d) {
        a[d * n + i] = p.getR()[d];
        a[(3 + d) * n + i] = p.getV()[d];
      }
      id[i] = (int64_t)p.getID();
    }
    if (!_periodic && (no != _uploadedOwned || id != _uploadedIds)) _structureChanged = true;
    if (_structureChanged) {
      std::vector<int32_t> type(n);
      for (size_t i = 0; i < n; ++i) type[i] = (int32_t)at(i).getTypeId();
      detail::check(_ctx,
                    ljg_upload(_ctx, (int64_t)n, &a[0], &a[n], &a[2 * n], &a[3 * n], &a[4 * n], &a[5 * n],
                               type.data(), id.data()),
                    "AutoPas: ljg_upload");
      if (!_periodic) detail::check(_ctx, ljg_set_owned(_ctx, (int64_t)no), "AutoPas: ljg_set_owned");
      _uploadedOwned = no;
      _uploadedIds = std::move(id);
      _structureChanged = false;
      _rebuildDue = true;
    } else {
      int32_t moved = 0;
      detail::check(_ctx,
                    ljg_update_positions(_ctx, &a[0], &a[n], &a[2 * n], &a[3 * n], &a[4 * n], &a[5 * n], &moved),
                    "AutoPas: ljg_update_positions");
      _rebuildDue = _rebuildDue || moved != 0;
    }
  }

  /* Device -> host particles: positions (wrapped at rebuilds in a periodic box), and either the pair
   * forces added to the owned particles' forces (computeInteractions) or the whole state (step). */
  void pullState(bool addForces) {
    const size_t no = _particles.size(), n = no + _halo.size();
    std::vector<double> a(9 * n);
    detail::check(_ctx,
                  ljg_download(_ctx, &a[0], &a[n], &a[2 * n], addForces ? nullptr : &a[3 * n],
                               addForces ? nullptr : &a[4 * n], addForces ? nullptr : &a[5 * n], &a[6 * n],
                               &a[7 * n], &a[8 * n]),
                  "AutoPas: ljg_download");
    for (size_t i = 0; i < no; ++i) {
      auto &p = _particles[i];
      p.setR({a[i], a[n + i], a[2 * n + i]});
      const std::array<double, 3> f{a[6 * n + i], a[7 * n + i], a[8 * n + i]};
      if (addForces) {
        p.addF(f);
      } else {
        p.setV({a[3 * n + i], a[4 * n + i], a[5 * n + i]});
        p.setF(f);
      }
    }
  }

  std::array<double, 3> _boxMin{0., 0., 0.}, _boxMax{10., 10., 10.};
  double _cutoff = 1.;
  double _skin = 0.2;
  unsigned int _rebuildFrequency = 10;
  unsigned int _clusterSize = 4;
  int32_t _kernel = -1;
  int32_t _device = 0;
  ljg_ctx *_ctx = nullptr;
  bool _periodic = true;
  std::vector<Particle_T> _particles;
  std::vector<Particle_T> _halo;       // open box: halo particles since the last updateContainer()
  size_t _uploadedOwned = 0;           // open box: owned count and ids (owned, then halo) of the last upload
  std::vector<int64_t> _uploadedIds;
  bool _structureChanged = true;
  bool _rebuildDue = true;
  unsigned int _stepsSinceRebuild = 0;
  size_t _numRebuilds = 0;
  size_t _iterations = 0;
  std::set<options::ContainerOption> _allowedContainers{options::ContainerOption::gpuLinkedCells,
                                                         options::ContainerOption::gpuVerletLists,
                                                         options::ContainerOption::gpuClusterLists};
  std::set<options::TraversalOption> _allowedTraversals{
      options::TraversalOption::gpu_lc_full_shell, options::TraversalOption::gpu_verlet_lists,
      options::TraversalOption::gpu_cluster_tiles, options::TraversalOption::gpu_cluster_jrows};
  std::set<options::Newton3Option> _allowedNewton3{options::Newton3Option::disabled, options::Newton3Option::enabled};
  std::set<options::PrecisionOption> _allowedPrecisions{options::PrecisionOption::fp32pair};
  std::set<int> _allowedClusterSizes{4, 8};
  std::set<int> _allowedTileGroups{0};
  unsigned int _tuningInterval = 5000;  // AutoTunerInfo.h
  unsigned int _maxSamples = 3;
  std::unique_ptr<AutoTuner> _tuner;
  Configuration _applied{};
  bool _haveApplied = false;
  int32_t _shift = 0;
  std::vector<double> _eps, _sig, _mass;
};

}  // namespace autopas_amd

#endif
