/* CPU oracle for the Coulomb-explosion hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Nothing in the product links, loads or calls this library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and there only
 * as the checker / the timed CPU baseline.
 *
 * What it restates (reference paths relative to the reference repository root):
 *   ceo_accelerations_* / ceo_potential_energy_* / ceo_kinetic_energy_f64
 *       src/system.cpp:5-49 (scalar force, PE, KE)
 *   ceo_sample_uniform_sphere
 *       src/sampler.cpp:7-57 (mt19937_64 + uniform_real_distribution, cube
 *       rejection, sequential min-separation placement)
 *   ceo_scalar_run_*
 *       src/integrators.cpp:59-250 (DormandPrince45) driven by
 *       src/driver.cpp:7-46 (run_to_convergence), plus the two extensions the
 *       engine adds: the momentum-change / t_max stop rule with FP64 time and a
 *       fixed initial step.
 *   ceo_batched_run_*
 *       include/coulomb/batched_integrator.hpp:62-433 (BatchedRK45<T>: shared-dt
 *       lockstep over K lanes, FSAL, worst-active-lane accept, latch, double
 *       finalize) on include/coulomb/batched_force.hpp:31-110.
 *   ceo_lane_run_*
 *       the per-lane formulation the GPU kernel executes (same DP5(4) method,
 *       position stages written in Runge-Kutta-Nystrom form so only the seven
 *       acceleration stages are stored), restated on the CPU.
 *   ceo_philox4x32_10 / ceo_device_sample
 *       the engine's own counter-based sampler (no reference counterpart).
 *   ceo_cpu_baseline
 *       the lockstep BatchedRK45<float> restatement at the host SIMD width,
 *       one instance per thread on disjoint sim ranges (the CPU baseline).
 */
#ifndef CE_ORACLE_H
#define CE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CEO_STOP_PE_FRACTION = 0, CEO_STOP_MOMENTUM_OR_TMAX = 1 };
enum { CEO_H0_HAIRER = 0, CEO_H0_FIXED = 1 };
enum {
  CEO_OK_PE = 0,
  CEO_OK_MOMENTUM = 1,
  CEO_OK_TMAX = 2,
  CEO_NOT_CONVERGED = 3,
  CEO_REJECT_LIMIT = 4,
  CEO_NONFINITE = 5,
  CEO_SAMPLER_FAIL = 6
};

typedef struct {
  double rtol, atol;
  double pe_stop_fraction;
  double max_dt;
  double t_max;
  double momentum_tol;
  double h0;
  uint64_t max_steps;
  int32_t stop_rule;
  int32_t redistribute;
  int32_t h0_mode;
  int32_t max_rejections;
} ceo_config;

typedef struct {
  int32_t n;
  int32_t pad_;
  const double* masses;
  const double* charges;
  double coulomb_k;
} ceo_molecule;

typedef struct {
  double* momenta;     /* [sim][atom][3] post-redistribution m*s*v */
  uint32_t* steps;     /* accepted steps (batched: batch iteration at latch) */
  uint32_t* attempts;  /* attempted steps, may be NULL */
  uint8_t* status;     /* CEO_OK_* / failure code */
  double* t_final;     /* may be NULL */
  double* energy_err;  /* (KE+PE-E0)/E0 at the stop, before redistribution; may be NULL */
  double* trace_h;     /* [sim][trace_cap] accepted step sizes; may be NULL */
  uint32_t trace_cap;
  uint32_t pad_;
} ceo_results;

/* ---- scalar physics (src/system.cpp) ---- */
void ceo_accelerations_f64(const ceo_molecule* mol, const double* pos, double* acc);
void ceo_accelerations_f32(const ceo_molecule* mol, const float* pos, float* acc);
double ceo_potential_energy_f64(const ceo_molecule* mol, const double* pos);
float ceo_potential_energy_f32(const ceo_molecule* mol, const float* pos);
double ceo_kinetic_energy_f64(const ceo_molecule* mol, const double* vel);

/* ---- sampler (src/sampler.cpp); one stream over n_sims consecutive draws.
 * Returns 0, -1 for bad options (invalid_argument), -2 when an atom cannot be
 * placed (runtime_error); sims before the failing one are filled. */
int ceo_sample_uniform_sphere(uint64_t seed, int32_t n_atoms, double radius, double min_sep,
                              uint64_t max_attempts, size_t n_sims, double* pos_out);

/* ---- scalar DP45 + driver, one independent run per sim ---- */
int ceo_scalar_run_f64(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                       size_t n_sims, ceo_results* out);
int ceo_scalar_run_f32(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                       size_t n_sims, ceo_results* out);

/* ---- lockstep batched DP5(4), lanes in {1,4,8,16}; sims padded per batch by
 * repeating the batch's last geometry (padding lanes are not reported). ---- */
int ceo_batched_run_f32(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                        size_t n_sims, int32_t lanes, ceo_results* out, uint64_t* batch_steps);
int ceo_batched_run_f64(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                        size_t n_sims, int32_t lanes, ceo_results* out, uint64_t* batch_steps);

/* ---- batched::accelerations / potential_energy<float> on the reference's SoA
 * layout p[i*k + lane] (batched_force.hpp:31-110), any k ---- */
void ceo_soa_accelerations_f32(size_t n, size_t k, const float* px, const float* py,
                               const float* pz, const float* pair_const, const float* inv_mass,
                               float* ax, float* ay, float* az);
void ceo_soa_potential_energy_f32(size_t n, size_t k, const float* px, const float* py,
                                  const float* pz, const float* pair_const, float* pe);

/* ---- per-lane formulation executed by the GPU kernel ----
 * ceo_lane_run_f32 takes the kernel's uniform form when every pair constant
 * k q_i q_j is the same c (finite, non-zero): forces summed with constant 1,
 * 1/m replaced by (float)(c / m), PE = c * sum (for c = 1 the same bits as the
 * table form). */
int ceo_lane_run_f32(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                     size_t n_sims, ceo_results* out);
int ceo_lane_run_f64(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                     size_t n_sims, ceo_results* out);
/* The same with per-sim charges charges[sim][atom] (the kernel's per-sim-charge
 * path: kq_i = (T)(k q_i), pair constant kq_i * q_j). */
int ceo_lane_run_q_f32(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                       const double* charges, size_t n_sims, ceo_results* out);
int ceo_lane_run_q_f64(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                       const double* charges, size_t n_sims, ceo_results* out);
/* The per-lane form stopped at t_cut and continued (momentum-or-t_max rule):
 * a sim that reaches t_cut by the time limit continues to cfg->t_max from its
 * state there, with the step size proposed after the step that landed on
 * t_cut -- the engine's saved-state resume.  charges may be NULL (the
 * molecule's); precision 0 = f32, 1 = f64.  -1 for t_cut <= 0 or another rule. */
int ceo_lane_run_cut(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                     const double* charges, size_t n_sims, double t_cut, int32_t precision,
                     ceo_results* out);
/* The derived coefficient tables of the per-lane form, in double:
 * c[7], aa[7*7] (A*A), b[7], e[7], ea[7] (E*A).  Written for inspection. */
void ceo_lane_coefficients(double* c, double* aa, double* b, double* e, double* ea);

/* ---- counter-based sampler ---- */
void ceo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Returns the number of sims whose placement failed. pos_out [sim][atom][3]. */
size_t ceo_device_sample(uint64_t seed, uint64_t first_index, size_t n_sims, int32_t n_atoms,
                         double radius, double min_sep, uint32_t max_attempts, double* pos_out,
                         uint8_t* status_out);

/* With per-sim charges uniform in {1..charge_levels} (q_out [sim][atom]), sums
 * capped by max_total (> 0). */
size_t ceo_device_sample_q(uint64_t seed, uint64_t first_index, size_t n_sims, int32_t n_atoms,
                           double radius, double min_sep, uint32_t max_attempts,
                           int32_t charge_levels, int32_t max_total, double* pos_out,
                           double* q_out, uint8_t* status_out);

/* ---- CPU baseline: lockstep f32 at the host SIMD width, n_threads instances.
 * Runs `reps` timed passes over the sims and reports the median wall time.
 * Returns the lane width used (0 on error). ---- */
int ceo_cpu_baseline(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                     size_t n_sims, int32_t n_threads, int32_t reps, double* median_seconds,
                     uint64_t* total_batch_steps, double* checksum);
const char* ceo_cpu_baseline_isa(void);
/* The baseline's own momenta (n_sims a multiple of its lane width), for
 * checking it against ceo_batched_run_f32.  Returns the lane width or 0. */
int ceo_cpu_baseline_momenta(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                             size_t n_sims, double* mom_out);

#ifdef __cplusplus
}
#endif

#endif /* CE_ORACLE_H */
