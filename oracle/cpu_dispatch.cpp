// Picks the widest CPU-baseline build the host supports -- TEST INFRASTRUCTURE ONLY.
#include "ce_oracle.h"

extern "C" int ceo_cpu_baseline_avx512(const ceo_molecule*, const ceo_config*, const double*,
                                       size_t, int32_t, int32_t, double*, uint64_t*, double*);
extern "C" int ceo_cpu_baseline_avx2(const ceo_molecule*, const ceo_config*, const double*,
                                     size_t, int32_t, int32_t, double*, uint64_t*, double*);
extern "C" int ceo_cpu_baseline_sse2(const ceo_molecule*, const ceo_config*, const double*,
                                     size_t, int32_t, int32_t, double*, uint64_t*, double*);

extern "C" int ceo_cpu_baseline_momenta_avx512(const ceo_molecule*, const ceo_config*,
                                               const double*, size_t, double*);
extern "C" int ceo_cpu_baseline_momenta_avx2(const ceo_molecule*, const ceo_config*,
                                             const double*, size_t, double*);
extern "C" int ceo_cpu_baseline_momenta_sse2(const ceo_molecule*, const ceo_config*,
                                             const double*, size_t, double*);

extern "C" int ceo_cpu_baseline_momenta(const ceo_molecule* mol, const ceo_config* cfg,
                                        const double* pos, size_t n_sims, double* mom) {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f"))
    return ceo_cpu_baseline_momenta_avx512(mol, cfg, pos, n_sims, mom);
  if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))
    return ceo_cpu_baseline_momenta_avx2(mol, cfg, pos, n_sims, mom);
  return ceo_cpu_baseline_momenta_sse2(mol, cfg, pos, n_sims, mom);
}

extern "C" const char* ceo_cpu_baseline_isa(void) {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f")) return "avx512 (16 f32 lanes)";
  if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return "avx2 (8 f32 lanes)";
  return "sse2 (4 f32 lanes)";
}

extern "C" int ceo_cpu_baseline(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                                size_t n_sims, int32_t n_threads, int32_t reps,
                                double* median_seconds, uint64_t* total_batch_steps,
                                double* checksum) {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f"))
    return ceo_cpu_baseline_avx512(mol, cfg, pos, n_sims, n_threads, reps, median_seconds,
                                   total_batch_steps, checksum);
  if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))
    return ceo_cpu_baseline_avx2(mol, cfg, pos, n_sims, n_threads, reps, median_seconds,
                                 total_batch_steps, checksum);
  return ceo_cpu_baseline_sse2(mol, cfg, pos, n_sims, n_threads, reps, median_seconds,
                               total_batch_steps, checksum);
}
