// Driver for the C++ host API (include/autopas_gpu.hpp), run by tests/test_cpp_api.py.
//
//   autopas_api_test errors            API misuse raises AutoPasException (no GPU needed for most)
//   autopas_api_test nodevice          init() without a GPU throws AutoPasException
//   autopas_api_test forces <file>     computeInteractions on a seeded lattice; writes x y z fx fy fz per
//                                      particle and the globals, for comparison with the oracle
//   autopas_api_test rebuilds          rebuild count of computeInteractions calls (LogicHandler's rule)
//   autopas_api_test tune <file>       computeInteractions with the AutoTuner over the GPU configurations
//   autopas_api_test loop <file>       md-flexible's step sequence through the API (host integrator,
//                                      TimeDiscretization.cpp:17-165) next to AutoPas::step on the device;
//                                      writes both end states
//   autopas_api_test halo <file> <k>   the same host steps on an open box with md-flexible's periodic
//                                      boundary made by the caller (leaving particles wrapped and re-added,
//                                      periodic images added as halo particles) next to the periodic box;
//                                      writes both end states and globals
//
// Every mode prints "ok <mode>" on success and exits 0; a failed check prints the reason and exits 1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "autopas_gpu.hpp"

using autopas_amd::AutoPas;
using autopas_amd::AutoPasException;
using autopas_amd::FunctorN3Modes;
using autopas_amd::MoleculeLJ;
using autopas_amd::ParticlePropertiesLibrary;

namespace {

constexpr double kCutoff = 2.5, kSkin = 0.3;

int fail(const std::string &why) {
  std::printf("FAIL %s\n", why.c_str());
  return 1;
}

template <class F>
bool throws(F &&fn, const char *needle) {
  try {
    fn();
  } catch (const AutoPasException &e) {
    return needle == nullptr || std::strstr(e.what(), needle) != nullptr;
  }
  return false;
}

struct Lattice {
  double L;
  std::vector<double> x, y, z, vx, vy, vz;
  std::vector<int32_t> type;
};

Lattice lattice(int n_side, double density, uint64_t seed, int ntypes) {
  Lattice s;
  const double a = std::pow(density, -1.0 / 3.0);
  s.L = n_side * a;
  const int64_t n = (int64_t)n_side * n_side * n_side;
  for (auto *v : {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz}) v->resize((size_t)n);
  s.type.resize((size_t)n);
  const double lo[3] = {0., 0., 0.};
  if (ljg_generate_lattice(seed, n, n_side, n_side, n_side, a, lo, 0.05, 0.9, ntypes, s.x.data(), s.y.data(),
                           s.z.data(), s.vx.data(), s.vy.data(), s.vz.data(), s.type.data()) != LJG_OK)
    throw AutoPasException("ljg_generate_lattice failed");
  return s;
}

void setup(AutoPas<MoleculeLJ> &ap, double L, int32_t kernel) {
  ap.setBoxMin({0., 0., 0.});
  ap.setBoxMax({L, L, L});
  ap.setCutoff(kCutoff);
  ap.setVerletSkin(kSkin);
  ap.setVerletRebuildFrequency(10);
  ap.setVerletClusterSize(4);
  ap.setKernel(kernel);
  ap.init();
}

void add_all(AutoPas<MoleculeLJ> &ap, const Lattice &s, bool types) {
  for (size_t i = 0; i < s.x.size(); ++i)
    ap.addParticle(MoleculeLJ({s.x[i], s.y[i], s.z[i]}, {s.vx[i], s.vy[i], s.vz[i]}, i,
                              types ? (unsigned long)s.type[i] : 0ul));
}

void write_state(FILE *f, const char *tag, AutoPas<MoleculeLJ> &ap) {
  std::fprintf(f, "%s %zu\n", tag, ap.getNumberOfParticles());
  for (auto &p : ap)
    std::fprintf(f, "%zu %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p.getID(), p.getR()[0],
                 p.getR()[1], p.getR()[2], p.getV()[0], p.getV()[1], p.getV()[2], p.getF()[0], p.getF()[1],
                 p.getF()[2]);
}

int mode_errors() {
  AutoPas<MoleculeLJ> ap;
  autopas_amd::LJFunctor<MoleculeLJ> f(kCutoff);
  if (!throws([&] { ap.computeInteractions(&f); }, "init() was not called"))
    return fail("computeInteractions before init");
  if (!throws([&] { f.getPotentialEnergy(); }, "calculateGlobals is false")) return fail("globals off");
  autopas_amd::LJFunctor<MoleculeLJ, true, false, FunctorN3Modes::Both, true> fg(kCutoff);
  if (!throws([&] { fg.getVirial(); }, "endTraversal was not called")) return fail("virial before traversal");
  MoleculeLJ a({0., 0., 0.}, {0., 0., 0.}, 0), b({1., 0., 0.}, {0., 0., 0.}, 1);
  if (!throws([&] { f.AoSFunctor(a, b, true); }, "not implemented")) return fail("AoSFunctor");
  if (!throws([&] { f.SoAFunctorPair(0, 0, true); }, "not implemented")) return fail("SoAFunctorPair");
  ap.setBoxMin({0., 0., 0.});
  ap.setBoxMax({10., 10., 10.});
  if (!throws([&] { ap.addParticle(MoleculeLJ({10., 1., 1.}, {0., 0., 0.}, 0)); }, "not in the bounding box"))
    return fail("particle outside the box");
  ap.addParticle(MoleculeLJ({1., 1., 1.}, {0., 0., 0.}, 0));
  if (ap.getNumberOfParticles() != 1) return fail("particle count");
  ParticlePropertiesLibrary ppl(kCutoff);
  if (!throws([&] { ppl.addSiteType(1, 1.); }, "consecutively")) return fail("site ids must be consecutive");
  ppl.addSiteType(0, 1.);
  ppl.addLJParametersToSite(0, 1., 1.);
  ppl.addSiteType(1, 2.);
  ppl.addLJParametersToSite(1, 1.21, 1.1);
  if (!throws([&] { ppl.getMixing24Epsilon(0, 1); }, "calculateMixingCoefficients")) return fail("mixing before calc");
  ppl.calculateMixingCoefficients();
  const double e01 = 24 * std::sqrt(1.21), s01 = 1.05 * 1.05;
  if (std::fabs(ppl.getMixing24Epsilon(0, 1) - e01) > 1e-15 * e01) return fail("mixed epsilon");
  if (std::fabs(ppl.getMixingSigmaSquared(1, 0) - s01) > 1e-15 * s01) return fail("mixed sigma");
  std::printf("ok errors\n");
  return 0;
}

int mode_nodevice() {
  AutoPas<MoleculeLJ> ap;
  ap.setBoxMin({0., 0., 0.});
  ap.setBoxMax({10., 10., 10.});
  ap.setCutoff(kCutoff);
  ap.setVerletSkin(kSkin);
  try {
    ap.init();
  } catch (const AutoPasException &e) {
    if (std::strstr(e.what(), "no HIP device") == nullptr) return fail(std::string("unexpected message: ") + e.what());
    std::printf("ok nodevice\n");
    return 0;
  }
  std::printf("ok nodevice (a device is present)\n");
  return 0;
}

// Forces and globals through computeInteractions: a single-type functor (setParticleProperties) on the
// j-row kernel and a two-type mixing functor on the cluster kernel, on the same lattice.
int mode_forces(const char *path) {
  FILE *out = std::fopen(path, "w");
  if (!out) return fail("cannot open output");
  const Lattice s = lattice(8, 0.75, 3, 2);
  {
    AutoPas<MoleculeLJ> ap;
    setup(ap, s.L, -1);  // default kernel (the j-row kernel for this box)
    add_all(ap, s, false);
    autopas_amd::LJFunctor<MoleculeLJ, true, false, FunctorN3Modes::Both, true, true> f(kCutoff);
    f.setParticleProperties(24., 1.);
    ap.computeInteractions(&f);
    std::fprintf(out, "globals %.17g %.17g %.6f %zu\n", f.getPotentialEnergy(), f.getVirial(), f.getHitRate(),
                 f.getNumFLOPs());
    write_state(out, "single", ap);
  }
  {
    AutoPas<MoleculeLJ> ap;
    setup(ap, s.L, LJG_KERNEL_CLUSTER_MIXED);
    add_all(ap, s, true);
    ParticlePropertiesLibrary ppl(kCutoff);
    ppl.addSiteType(0, 1.);
    ppl.addLJParametersToSite(0, 1., 1.);
    ppl.addSiteType(1, 1.);
    ppl.addLJParametersToSite(1, 1.2, 1.05);
    ppl.calculateMixingCoefficients();
    autopas_amd::LJFunctor<MoleculeLJ, true, true, FunctorN3Modes::Both, true> f(kCutoff, ppl);
    ap.computeInteractions(&f);
    std::fprintf(out, "globals %.17g %.17g 0 0\n", f.getPotentialEnergy(), f.getVirial());
    write_state(out, "mixing", ap);
  }
  std::fclose(out);
  std::printf("ok forces\n");
  return 0;
}

// md-flexible's step: calculatePositionsAndResetForces, updateContainer, computeInteractions,
// calculateVelocities (TimeDiscretization.cpp:17-69,152-165), with the factors formed as there.
template <class Functor>
void host_step(AutoPas<MoleculeLJ> &ap, Functor &f, double dt) {
  for (auto &p : ap) {
    const double ff = dt * dt / (2 * 1.0);
    const auto &v = p.getV();
    const auto &F = p.getF();
    p.addR({v[0] * dt + F[0] * ff, v[1] * dt + F[1] * ff, v[2] * dt + F[2] * ff});
    p.setOldF(F);
    p.setF({0., 0., 0.});
  }
  (void)ap.updateContainer();
  ap.computeInteractions(&f);
  for (auto &p : ap) {
    const double kf = dt / (2 * 1.0);
    const auto &F = p.getF();
    const auto &oF = p.getOldF();
    p.addV({(F[0] + oF[0]) * kf, (F[1] + oF[1]) * kf, (F[2] + oF[2]) * kf});
  }
}

int mode_loop(const char *path, int32_t kernel) {
  FILE *out = std::fopen(path, "w");
  if (!out) return fail("cannot open output");
  const Lattice s = lattice(8, 0.75, 5, 1);
  const double dt = 0.002;
  const int steps = 25;
  // the host loop starts from zero force, as the device loop does after the upload
  AutoPas<MoleculeLJ> host;
  setup(host, s.L, kernel);
  add_all(host, s, false);
  autopas_amd::LJFunctor<MoleculeLJ, true> f(kCutoff);
  f.setParticleProperties(24., 1.);
  for (int k = 0; k < steps; ++k) host_step(host, f, dt);
  write_state(out, "host", host);
  AutoPas<MoleculeLJ> dev;
  setup(dev, s.L, kernel);
  add_all(dev, s, false);
  const ljg_run_stats rs = dev.step(steps, dt);
  std::fprintf(out, "stats %lld %lld\n", (long long)rs.steps, (long long)rs.rebuilds);
  write_state(out, "device", dev);
  std::fclose(out);
  std::printf("ok loop\n");
  return 0;
}

// md-flexible's periodic boundary on an open box (RegularGridDecomposition.cpp:515-587): after the drift,
// updateContainer returns the owned particles that left the box, which are wrapped back and re-added, and
// every periodic image within rc + skin of the box is added as a halo particle; then computeInteractions
// and the kick on the owned particles.
template <class Functor>
void host_step_open(AutoPas<MoleculeLJ> &ap, Functor &f, double dt, double L) {
  for (auto &p : ap) {
    const double ff = dt * dt / (2 * 1.0);
    const auto &v = p.getV();
    const auto &F = p.getF();
    p.addR({v[0] * dt + F[0] * ff, v[1] * dt + F[1] * ff, v[2] * dt + F[2] * ff});
    p.setOldF(F);
    p.setF({0., 0., 0.});
  }
  for (auto p : ap.updateContainer()) {
    std::array<double, 3> r = p.getR();
    for (int d = 0; d < 3; ++d) {
      if (r[d] < 0.) r[d] += L;
      else if (r[d] >= L) r[d] -= L;
    }
    p.setR(r);
    ap.addParticle(p);
  }
  const double w = kCutoff + kSkin;
  std::vector<MoleculeLJ> halo;
  for (auto &p : ap)
    for (int sx = -1; sx <= 1; ++sx)
      for (int sy = -1; sy <= 1; ++sy)
        for (int sz = -1; sz <= 1; ++sz) {
          if (sx == 0 && sy == 0 && sz == 0) continue;
          const std::array<double, 3> r{p.getR()[0] + sx * L, p.getR()[1] + sy * L, p.getR()[2] + sz * L};
          bool near = true;
          for (int d = 0; d < 3; ++d) near = near && r[d] >= -w && r[d] <= L + w;
          if (!near) continue;
          MoleculeLJ h = p;
          h.setR(r);
          halo.push_back(h);
        }
  for (auto &h : halo) ap.addHaloParticle(h);
  ap.computeInteractions(&f);
  for (auto &p : ap) {
    const double kf = dt / (2 * 1.0);
    const auto &F = p.getF();
    const auto &oF = p.getOldF();
    p.addV({(F[0] + oF[0]) * kf, (F[1] + oF[1]) * kf, (F[2] + oF[2]) * kf});
  }
}

int mode_halo(const char *path, int32_t kernel) {
  FILE *out = std::fopen(path, "w");
  if (!out) return fail("cannot open output");
  const Lattice s = lattice(8, 0.75, 9, 1);
  const double dt = 0.002;
  const int steps = 30;
  autopas_amd::LJFunctor<MoleculeLJ, true, false, FunctorN3Modes::Both, true> f(kCutoff);
  f.setParticleProperties(24., 1.);
  AutoPas<MoleculeLJ> open;
  open.setPeriodicBoundaries(false);
  setup(open, s.L, kernel);
  add_all(open, s, false);
  // halo particles belong to open boxes only, and must lie outside the box
  if (!throws([&] { open.addHaloParticle(MoleculeLJ({1., 1., 1.}, {0., 0., 0.}, 0, 0)); }, "not outside the box"))
    return fail("a halo particle inside the box was accepted");
  size_t halo = 0;
  for (int k = 0; k < steps; ++k) {
    host_step_open(open, f, dt, s.L);
    halo += open.getNumberOfParticles(autopas_amd::IteratorBehavior::halo);
  }
  const double eo = f.getPotentialEnergy(), vo = f.getVirial();
  if (open.getNumberOfParticles() != s.x.size()) return fail("owned particles lost on the open box");
  std::fprintf(out, "globals %.17g %.17g %zu %zu\n", eo, vo, halo / steps, open.getNumRebuilds());
  write_state(out, "open", open);
  AutoPas<MoleculeLJ> per;
  setup(per, s.L, kernel);
  add_all(per, s, false);
  if (!throws([&] { per.addHaloParticle(MoleculeLJ({-1., 1., 1.}, {0., 0., 0.}, 0, 0)); }, "periodic"))
    return fail("a halo particle was accepted by a periodic box");
  for (int k = 0; k < steps; ++k) host_step(per, f, dt);
  std::fprintf(out, "globals %.17g %.17g 0 %zu\n", f.getPotentialEnergy(), f.getVirial(), per.getNumRebuilds());
  write_state(out, "periodic", per);
  std::fclose(out);
  std::printf("ok halo\n");
  return 0;
}

// LogicHandler's rebuild rule through computeInteractions (LogicHandler.h:987-1015): K calls that move
// every particle by much less than skin/2 rebuild on the first call and then every rebuild frequency
// calls; one particle moved by skin/2 since the last rebuild forces a rebuild at the next call.
int mode_rebuilds() {
  const Lattice s = lattice(8, 0.75, 7, 1);
  AutoPas<MoleculeLJ> ap;
  setup(ap, s.L, LJG_KERNEL_CLUSTER_MIXED);  // rebuild frequency 10
  add_all(ap, s, false);
  autopas_amd::LJFunctor<MoleculeLJ, true> f(kCutoff);
  f.setParticleProperties(24., 1.);
  const int calls = 25;
  for (int k = 0; k < calls; ++k) {
    for (auto &p : ap) p.addR({1e-4, -1e-4, 0.});  // 25 calls: 3.5e-3 in all, below skin/2 = 0.15
    ap.computeInteractions(&f);
  }
  if (ap.getNumRebuilds() != 3)
    return fail("rebuilds after 25 calls: " + std::to_string(ap.getNumRebuilds()) + ", expected 3 (calls 1, 11, 21)");
  ap.begin()->addR({0.16, 0., 0.});  // > skin/2 from where it was at the last rebuild
  ap.computeInteractions(&f);
  if (ap.getNumRebuilds() != 4) return fail("a move of skin/2 did not trigger a rebuild");
  ap.computeInteractions(&f);
  if (ap.getNumRebuilds() != 4) return fail("the rebuild after the move did not reset the trigger");
  const double mean = ap.getMeanRebuildFrequency();
  if (std::fabs(mean - 27. / 4.) > 1e-12) return fail("mean rebuild frequency " + std::to_string(mean));
  std::printf("ok rebuilds\n");
  return 0;
}

// computeInteractions with LJG_KERNEL_AUTO: the AutoTuner runs one tuning phase over the GPU
// configurations (cartesian product of the allowed options), maxSamples iterations each, then keeps the
// configuration with the lowest effective time.  Writes the search space, the evidence, the chosen
// configuration and the final state (positions and forces of the last call) for comparison with the oracle.
int mode_tune(const char *path) {
  FILE *out = std::fopen(path, "w");
  if (!out) return fail("cannot open output");
  namespace o = autopas_amd::options;
  const Lattice s = lattice(12, 0.75, 9, 1);  // L = 13.2: every kernel applies (> 3 (rc + skin))
  AutoPas<MoleculeLJ> ap;
  setup(ap, s.L, LJG_KERNEL_AUTO);
  ap.setAllowedContainers({o::ContainerOption::gpuLinkedCells, o::ContainerOption::gpuVerletLists,
                           o::ContainerOption::gpuClusterLists});
  ap.setAllowedPrecisions({o::PrecisionOption::fp64, o::PrecisionOption::fp32pair});
  ap.setAllowedNewton3Options({o::Newton3Option::disabled, o::Newton3Option::enabled});
  ap.setAllowedClusterSizes({4, 8});
  ap.setMaxSamples(3);
  ap.setTuningInterval(1000);
  add_all(ap, s, false);
  autopas_amd::LJFunctor<MoleculeLJ, true, false, FunctorN3Modes::Both, true> f(kCutoff);
  f.setParticleProperties(24., 1.);
  int tuningCalls = 0, calls = 0;
  std::set<std::string> seen;
  for (;; ++calls) {
    for (auto &p : ap) p.setF({0., 0., 0.});
    const bool tuning = ap.computeInteractions(&f);
    seen.insert(ap.getCurrentConfig().toString());
    if (!tuning) break;
    ++tuningCalls;
    if (calls > 200) return fail("the tuning phase did not end");
  }
  const auto *tuner = ap.getAutoTuner();
  if (!tuner || tuner->inTuningPhase()) return fail("no tuner or still tuning");
  const size_t space = tuner->getSearchSpace().size();
  std::fprintf(out, "# space %zu calls %d tuning_calls %d seen %zu\n", space, calls + 1, tuningCalls, seen.size());
  long best = -1;
  for (const auto &kv : tuner->getEvidence()) {
    std::fprintf(out, "# evidence %s %ld\n", kv.first.toString().c_str(), kv.second);
    if (best < 0 || kv.second < best) best = kv.second;
  }
  const auto &chosen = ap.getCurrentConfig();
  const auto it = tuner->getEvidence().find(chosen);
  if (it == tuner->getEvidence().end() || it->second != best) return fail("the chosen configuration is not the fastest");
  std::fprintf(out, "# chosen %s kernel %d\n", chosen.toString().c_str(), (int)chosen.kernel());
  std::fprintf(out, "globals %.17g %.17g 0 0\n", f.getPotentialEnergy(), f.getVirial());
  write_state(out, "final", ap);
  std::fclose(out);
  std::printf("ok tune\n");
  return 0;
}

}  // namespace

// FLOP model of the functor (LJFunctor.h:722-789) after a full-list and a Newton-3 kernel on the same
// lattice: one line "flops <kernel> <hit rate> <FLOPs>" each.
int mode_flops(const char *path) {
  FILE *out = std::fopen(path, "w");
  if (!out) return fail("cannot open output");
  const Lattice s = lattice(8, 0.75, 3, 1);
  for (int32_t kernel : {LJG_KERNEL_CLUSTER_FP64, LJG_KERNEL_CLUSTER_FP64_N3}) {
    AutoPas<MoleculeLJ> ap;
    setup(ap, s.L, kernel);
    add_all(ap, s, false);
    autopas_amd::LJFunctor<MoleculeLJ, true, false, FunctorN3Modes::Both, true, true> f(kCutoff);
    f.setParticleProperties(24., 1.);
    ap.computeInteractions(&f);
    ap.computeInteractions(&f);  // the counters start again with each traversal
    std::fprintf(out, "flops %d %.17g %zu\n", (int)kernel, f.getHitRate(), f.getNumFLOPs());
  }
  std::fclose(out);
  std::printf("ok flops\n");
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) return fail("usage: autopas_api_test errors|nodevice|forces <file>|loop <file> [kernel]");
  const std::string mode = argv[1];
  try {
    if (mode == "errors") return mode_errors();
    if (mode == "nodevice") return mode_nodevice();
    if (mode == "rebuilds") return mode_rebuilds();
    if (mode == "tune" && argc >= 3) return mode_tune(argv[2]);
    if (mode == "forces" && argc >= 3) return mode_forces(argv[2]);
    if (mode == "flops" && argc >= 3) return mode_flops(argv[2]);
    if (mode == "loop" && argc >= 3) return mode_loop(argv[2], argc >= 4 ? std::atoi(argv[3]) : -1);
    if (mode == "halo" && argc >= 3) return mode_halo(argv[2], argc >= 4 ? std::atoi(argv[3]) : -1);
  } catch (const AutoPasException &e) {
    return fail(std::string("AutoPasException: ") + e.what());
  }
  return fail("unknown mode " + mode);
}
