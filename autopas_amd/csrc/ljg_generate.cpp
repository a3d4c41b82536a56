// ljg_generate.cpp -- seeded synthetic LJ systems (host side).  Random numbers come from Philox4x32-10
// (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3", SC'11), a counter-based
// generator, so every coordinate is a pure function of (seed, particle, stream) and independent of
// threading.  Normal deviates use the Box-Muller transform.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/ljgpu.h"

namespace {

struct U4 {
  uint32_t v[4];
};

inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// Philox4x32 with 10 rounds: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 /
// 0xBB67AE85, as published.
U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, ctr.v[0], hi0, lo0);
    mulhilo(0xCD9E8D57u, ctr.v[2], hi1, lo1);
    ctr = U4{{hi1 ^ ctr.v[1] ^ k0, lo1, hi0 ^ ctr.v[3] ^ k1, lo0}};
  }
  return ctr;
}

// Two uniform doubles in [0, 1) with 53 random bits each, from one Philox block.
inline void uniform2(uint64_t seed, uint64_t index, uint32_t stream, double &u0, double &u1) {
  const U4 r = philox4x32_10(U4{{(uint32_t)index, (uint32_t)(index >> 32), stream, 0u}}, (uint32_t)seed,
                             (uint32_t)(seed >> 32));
  const uint64_t a = ((uint64_t)r.v[0] << 32) | r.v[1];
  const uint64_t b = ((uint64_t)r.v[2] << 32) | r.v[3];
  u0 = (double)(a >> 11) * 0x1.0p-53;
  u1 = (double)(b >> 11) * 0x1.0p-53;
}

inline double normal(uint64_t seed, uint64_t index, uint32_t stream) {
  double u0, u1;
  uniform2(seed, index, stream, u0, u1);
  const double r = std::sqrt(-2.0 * std::log(1.0 - u0));
  return r * std::cos(2.0 * 3.14159265358979323846 * u1);
}

// Maxwell-Boltzmann velocities at temperature T for mass 1 (streams 3-5), net momentum removed, rescaled to
// T exactly with 3n - 3 degrees of freedom.
void thermal_velocities(uint64_t seed, int64_t n, double temperature, double *vx, double *vy, double *vz) {
  const double sd = std::sqrt(temperature);
  double m[3] = {0., 0., 0.};
  for (int64_t s = 0; s < n; ++s) {
    vx[s] = sd * normal(seed, (uint64_t)s, 3u);
    vy[s] = sd * normal(seed, (uint64_t)s, 4u);
    vz[s] = sd * normal(seed, (uint64_t)s, 5u);
    m[0] += vx[s];
    m[1] += vy[s];
    m[2] += vz[s];
  }
  for (int d = 0; d < 3; ++d) m[d] /= (double)n;
  double ekin2 = 0.;
  for (int64_t s = 0; s < n; ++s) {
    vx[s] -= m[0];
    vy[s] -= m[1];
    vz[s] -= m[2];
    ekin2 += vx[s] * vx[s] + vy[s] * vy[s] + vz[s] * vz[s];
  }
  const double dof = n > 1 ? 3.0 * (double)n - 3.0 : 3.0;
  const double tcur = ekin2 / dof;
  const double scale = tcur > 0. ? std::sqrt(temperature / tcur) : 0.;
  for (int64_t s = 0; s < n; ++s) {
    vx[s] *= scale;
    vy[s] *= scale;
    vz[s] *= scale;
  }
}

// Uniform type in [0, ntypes) for particle s (stream 2).
inline int32_t uniform_type(uint64_t seed, int64_t s, int32_t ntypes) {
  if (ntypes <= 1) return 0;
  double t0, t1;
  uniform2(seed, (uint64_t)s, 2u, t0, t1);
  const int32_t t = (int32_t)(t0 * ntypes);
  return t < ntypes ? t : ntypes - 1;
}

}  // namespace

extern "C" ljg_status ljg_generate_lattice(uint64_t seed, int64_t n, int64_t nx, int64_t ny, int64_t nz, double a,
                                           const double box_min[3], double jitter, double temperature,
                                           int32_t ntypes, double *x, double *y, double *z, double *vx, double *vy,
                                           double *vz, int32_t *type) {
  if (n <= 0 || nx <= 0 || ny <= 0 || nz <= 0 || n > nx * ny * nz || !box_min || !x || !y || !z) return LJG_ERR_ARG;
  for (int64_t s = 0; s < n; ++s) {
    const int64_t ix = s % nx, iy = (s / nx) % ny, iz = s / (nx * ny);
    double j0, j1, j2, j3;
    uniform2(seed, (uint64_t)s, 0u, j0, j1);
    uniform2(seed, (uint64_t)s, 1u, j2, j3);
    x[s] = box_min[0] + (ix + 0.5) * a + (2.0 * j0 - 1.0) * jitter;
    y[s] = box_min[1] + (iy + 0.5) * a + (2.0 * j1 - 1.0) * jitter;
    z[s] = box_min[2] + (iz + 0.5) * a + (2.0 * j2 - 1.0) * jitter;
    if (type) type[s] = uniform_type(seed, s, ntypes);
  }
  if (vx && vy && vz) thermal_velocities(seed, n, temperature, vx, vy, vz);
  return LJG_OK;
}

// Two-phase system by random sequential placement: candidates are uniform in the slab (liquid, first
// n_liquid particles) or in its complement (gas, the rest); a candidate is kept when no earlier particle
// lies closer than dmin (periodic minimum image).  Earlier particles are found through a periodic grid of
// cells no smaller than dmin (head/next lists, 27 neighbour cells).  Candidate k uses Philox block
// (seed, k) on streams 6 and 7, so the system depends only on the seed.
extern "C" ljg_status ljg_generate_slab(uint64_t seed, const double box[3], double slab_lo, double slab_hi,
                                        int64_t n_liquid, int64_t n_gas, double dmin, int32_t ntypes,
                                        double temperature, int64_t max_attempts, double *x, double *y, double *z,
                                        double *vx, double *vy, double *vz, int32_t *type) {
  if (!box || !x || !y || !z || n_liquid < 0 || n_gas < 0 || n_liquid + n_gas <= 0 || !(dmin > 0.)) return LJG_ERR_ARG;
  if (!(slab_lo >= 0. && slab_hi > slab_lo && slab_hi <= box[2])) return LJG_ERR_ARG;
  if (n_gas > 0 && !(box[2] - (slab_hi - slab_lo) > 0.)) return LJG_ERR_ARG;
  int64_t nc[3];
  for (int d = 0; d < 3; ++d) {
    if (!(box[d] >= 3. * dmin)) return LJG_ERR_ARG;
    nc[d] = (int64_t)std::floor(box[d] / dmin);
  }
  const int64_t ncell = nc[0] * nc[1] * nc[2];
  const int64_t n = n_liquid + n_gas;
  std::vector<int64_t> head((size_t)ncell, -1), next((size_t)n, -1);
  const double d2min = dmin * dmin;
  auto cell_of = [&](double p, int d) {
    int64_t c = (int64_t)std::floor(p / box[d] * (double)nc[d]);
    return c < 0 ? 0 : (c >= nc[d] ? nc[d] - 1 : c);
  };
  uint64_t attempt = 0;
  const double thick = slab_hi - slab_lo, gas_len = box[2] - thick;
  for (int64_t s = 0; s < n; ++s) {
    for (;;) {
      if ((int64_t)attempt >= max_attempts) return LJG_ERR_CAPACITY;
      double u0, u1, u2, u3;
      uniform2(seed, attempt, 6u, u0, u1);
      uniform2(seed, attempt, 7u, u2, u3);
      ++attempt;
      const double px = u0 * box[0], py = u1 * box[1];
      double pz;
      if (s < n_liquid) {
        pz = slab_lo + u2 * thick;
      } else {
        const double g = u2 * gas_len;  // gas volume: [0, slab_lo) then [slab_hi, box_z)
        pz = g < slab_lo ? g : g + thick;
      }
      const int64_t cx = cell_of(px, 0), cy = cell_of(py, 1), cz = cell_of(pz, 2);
      bool ok = true;
      for (int oz = -1; oz <= 1 && ok; ++oz) {
        const int64_t jz = (cz + oz + nc[2]) % nc[2];
        for (int oy = -1; oy <= 1 && ok; ++oy) {
          const int64_t jy = (cy + oy + nc[1]) % nc[1];
          for (int ox = -1; ox <= 1 && ok; ++ox) {
            const int64_t jx = (cx + ox + nc[0]) % nc[0];
            for (int64_t q = head[(size_t)((jz * nc[1] + jy) * nc[0] + jx)]; q >= 0; q = next[(size_t)q]) {
              double dx = px - x[q], dy = py - y[q], dz = pz - z[q];
              dx -= box[0] * std::round(dx / box[0]);
              dy -= box[1] * std::round(dy / box[1]);
              dz -= box[2] * std::round(dz / box[2]);
              if (dx * dx + dy * dy + dz * dz < d2min) {
                ok = false;
                break;
              }
            }
          }
        }
      }
      if (!ok) continue;
      x[s] = px;
      y[s] = py;
      z[s] = pz;
      const size_t c = (size_t)((cz * nc[1] + cy) * nc[0] + cx);
      next[(size_t)s] = head[c];
      head[c] = s;
      break;
    }
    if (type) type[s] = uniform_type(seed, s, ntypes);
  }
  if (vx && vy && vz) thermal_velocities(seed, n, temperature, vx, vy, vz);
  return LJG_OK;
}
