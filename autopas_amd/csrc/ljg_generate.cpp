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
    if (type) {
      if (ntypes > 1) {
        double t0, t1;
        uniform2(seed, (uint64_t)s, 2u, t0, t1);
        int32_t t = (int32_t)(t0 * ntypes);
        type[s] = t < ntypes ? t : ntypes - 1;
      } else {
        type[s] = 0;
      }
    }
  }
  if (vx && vy && vz) {
    const double sd = std::sqrt(temperature);  // mass 1
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
  return LJG_OK;
}
