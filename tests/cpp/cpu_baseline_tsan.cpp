// The threaded CPU baseline (oracle/cpu_baseline.cpp, SSE2 build) under
// ThreadSanitizer: several timed passes over one sim set on 4 threads, each
// thread on its own range.  The summed batch steps must not depend on the
// thread count (every batch runs the same on any thread); a shared write between
// the workers is a data race TSan reports.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "ce_oracle.h"

extern "C" int ceo_cpu_baseline_sse2(const ceo_molecule*, const ceo_config*, const double*,
                                     size_t, int32_t, int32_t, double*, uint64_t*, double*);

int main() {
  const double masses[5] = {21874.66, 1837.15, 1837.15, 1837.15, 1837.15};
  const double charges[5] = {1, 1, 1, 1, 1};
  const ceo_molecule mol{5, 0, masses, charges, 1.0};
  ceo_config cfg{};
  cfg.rtol = 1e-4;
  cfg.atol = 1e-7;
  cfg.pe_stop_fraction = 1e-5;
  cfg.max_dt = 1e300;
  cfg.t_max = 1e300;
  cfg.max_steps = 1000000;
  cfg.stop_rule = 0;
  cfg.redistribute = 1;
  cfg.h0_mode = 0;
  cfg.max_rejections = 100;
  // CH4-like geometries: a tetrahedron around the carbon, each atom moved by up
  // to 0.3 bohr (the baseline does not sample; it integrates what it is given)
  const size_t n_sims = 4 * 4 * 12;
  const double tet[5][3] = {{0, 0, 0}, {1.2, 1.2, 1.2}, {-1.2, -1.2, 1.2}, {-1.2, 1.2, -1.2},
                            {1.2, -1.2, -1.2}};
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> jitter(-0.3, 0.3);
  std::vector<double> pos(n_sims * 15);
  for (size_t s = 0; s < n_sims; ++s)
    for (int i = 0; i < 5; ++i)
      for (int c = 0; c < 3; ++c) pos[(s * 5 + i) * 3 + c] = tet[i][c] + jitter(rng);
  double sec1 = 0, sec4 = 0, cs1 = 0, cs4 = 0;
  uint64_t st1 = 0, st4 = 0;
  const int k1 = ceo_cpu_baseline_sse2(&mol, &cfg, pos.data(), n_sims, 1, 1, &sec1, &st1, &cs1);
  const int k4 = ceo_cpu_baseline_sse2(&mol, &cfg, pos.data(), n_sims, 4, 3, &sec4, &st4, &cs4);
  if (k1 != 4 || k4 != 4 || st1 == 0 || st1 != st4 || !(std::fabs(cs4 - cs1) <= 1e-9 * std::fabs(cs1))) {
    std::fprintf(stderr, "baseline mismatch: lanes %d/%d steps %llu/%llu checksum %.17g/%.17g\n", k1,
                 k4, static_cast<unsigned long long>(st1), static_cast<unsigned long long>(st4),
                 cs1, cs4);
    return 1;
  }
  std::printf("cpu baseline ok: %zu sims, %llu batch steps on 1 and 4 threads\n", n_sims,
              static_cast<unsigned long long>(st4));
  return 0;
}
