// Host sampler with the reference's stream (include/coulomb/sampler.hpp:46-72,
// src/sampler.cpp:7-57): the seeded std::mt19937_64 / uniform_real_distribution
// pair the reference uses, so a seed gives the reference's geometries exactly.
// The engine's own device sampler (kernels.hpp, sample_sphere) is the fast path.
#include <cstdint>
#include <random>
#include <string>

#include "coulomb_explosion.h"

namespace ce {
int set_error(int code, const std::string& msg);
}

struct ce_host_sampler {
  double radius;
  double min_sep;
  uint64_t max_attempts;
  std::mt19937_64 rng;
};

extern "C" {

int ce_host_sampler_create(uint64_t seed, double radius, double min_separation,
                           uint64_t max_attempts, ce_host_sampler** out) {
  if (out == nullptr) return ce::set_error(CE_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!(radius > 0.0))
    return ce::set_error(CE_ERR_INVALID_ARGUMENT, "UniformSphereSampler: radius must be positive");
  if (!(min_separation >= 0.0))
    return ce::set_error(CE_ERR_INVALID_ARGUMENT,
                         "UniformSphereSampler: min_separation must be non-negative");
  *out = new ce_host_sampler{radius, min_separation, max_attempts, std::mt19937_64(seed)};
  return CE_OK;
}

int ce_host_sampler_reseed(ce_host_sampler* s, uint64_t seed) {
  if (s == nullptr) return ce::set_error(CE_ERR_INVALID_ARGUMENT, "sampler is NULL");
  s->rng.seed(seed);
  return CE_OK;
}

int ce_host_sampler_sample(ce_host_sampler* s, int32_t n_atoms, size_t n_sims, double* pos_out) {
  if (s == nullptr || (pos_out == nullptr && n_sims > 0))
    return ce::set_error(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_atoms < 0) return ce::set_error(CE_ERR_INVALID_ARGUMENT, "n_atoms < 0");
  const double R = s->radius, r2 = R * R, min2 = s->min_sep * s->min_sep;
  std::uniform_real_distribution<double> coord(-R, R);
  for (size_t k = 0; k < n_sims; ++k) {
    double* P = pos_out + k * static_cast<size_t>(n_atoms) * 3;
    for (int i = 0; i < n_atoms; ++i) {
      uint64_t fails = 0;
      for (;;) {
        double x, y, z;
        do {  // uniform in the ball by rejection in the bounding cube
          x = coord(s->rng);
          y = coord(s->rng);
          z = coord(s->rng);
        } while (!(x * x + y * y + z * z <= r2));
        bool clears = true;
        for (int j = 0; j < i && clears; ++j) {
          const double dx = x - P[3 * j], dy = y - P[3 * j + 1], dz = z - P[3 * j + 2];
          clears = !(dx * dx + dy * dy + dz * dz < min2);
        }
        if (clears) {
          P[3 * i] = x;
          P[3 * i + 1] = y;
          P[3 * i + 2] = z;
          break;
        }
        if (++fails >= s->max_attempts)
          return ce::set_error(
              CE_ERR_RUNTIME,
              "UniformSphereSampler: could not place an atom respecting min_separation; "
              "radius too small or min_separation too large for the atom count");
      }
    }
  }
  return CE_OK;
}

int ce_host_sampler_destroy(ce_host_sampler* s) {
  delete s;
  return CE_OK;
}

}  // extern "C"
