/* coulomb_explosion.h -- C-ABI of the MI355X (gfx950) Coulomb-explosion engine.
 *
 * This is the drop-in boundary for the reference's batched f32 path: the
 * reference's BatchedRK45<float> (include/coulomb/batched_integrator.hpp:62-191)
 * with its sampler (include/coulomb/sampler.hpp:46-72) and its batched kernels
 * (include/coulomb/batched_force.hpp:31-110).  The reference has no FFI; its
 * C++ interface maps onto these entry points as noted per function, and
 * include/coulomb_gpu.hpp wraps them back into that C++ interface.
 *
 * Conventions
 *   - Plain pointers and sizes only; every entry point returns int (CE_OK = 0,
 *     negative = error) and never throws.  ce_last_error() gives the message of
 *     the last failure on the calling thread.
 *   - CE_ERR_INVALID_ARGUMENT corresponds to the reference's
 *     std::invalid_argument, CE_ERR_RUNTIME to std::runtime_error.
 *   - An engine owns its device buffers and one HIP stream.  It is
 *     thread-compatible (one engine per host thread), like the reference
 *     instance.  Host pointers are host memory unless a name says _device.
 *   - Positions are atomic units (bohr), masses electron masses, time a.u.
 */
#ifndef COULOMB_EXPLOSION_H
#define COULOMB_EXPLOSION_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CE_MAX_ATOMS 10

enum ce_error {
  CE_OK = 0,
  CE_ERR_INVALID_ARGUMENT = -1,
  CE_ERR_RUNTIME = -2,
  CE_ERR_HIP = -3,
  CE_ERR_NO_DEVICE = -4
};

/* Termination rule.  CE_STOP_PE_FRACTION is the reference's
 * (driver.hpp:28-35, batched_integrator.hpp:149-169): stop once
 * PE <= pe_stop_fraction * PE0.  CE_STOP_MOMENTUM_OR_TMAX stops when every
 * atom's momentum changed by less than momentum_tol (relative, >= 0) over the
 * last accepted step, or when the FP64 time reaches t_max (> 0; the last step
 * is clamped to land on t_max, and the momentum test is not applied to that
 * shortened step: the sim ends with CE_SIM_OK_TMAX). */
enum ce_stop_rule { CE_STOP_PE_FRACTION = 0, CE_STOP_MOMENTUM_OR_TMAX = 1 };
/* Initial step: the reference's Hairer heuristic (integrators.cpp:197-250)
 * or a fixed h0. */
enum ce_h0_mode { CE_H0_HAIRER = 0, CE_H0_FIXED = 1 };
/* Arithmetic of the integration: f32 state (production) or the f64 path used
 * as the on-GPU accuracy oracle. */
enum ce_precision { CE_PRECISION_F32 = 0, CE_PRECISION_F64 = 1 };

/* Per-simulation status byte. */
enum ce_sim_status {
  CE_SIM_OK_PE = 0,         /* PE fraction reached */
  CE_SIM_OK_MOMENTUM = 1,   /* momentum settled */
  CE_SIM_OK_TMAX = 2,       /* reached t_max */
  CE_SIM_NOT_CONVERGED = 3, /* max_steps hit first (reference: converged=0) */
  CE_SIM_REJECT_LIMIT = 4,  /* > max_rejections in one step (reference throws) */
  CE_SIM_NONFINITE = 5,     /* error norm NaN through the rejection limit */
  CE_SIM_SAMPLER_FAIL = 6,  /* device sampler could not place the atoms */
  CE_SIM_CLOSURE = 7        /* momenta predicted by the learned closure */
};

/* Molecule: replaces the reference's Molecule (molecule.hpp:30-35) and the
 * BatchedRK45 constructor's pair_const / inv_mass set-up
 * (batched_integrator.hpp:71-106).  1 <= n_atoms <= CE_MAX_ATOMS. */
typedef struct {
  int32_t n_atoms;
  int32_t pad_;
  const double* masses;  /* n_atoms, electron masses */
  const double* charges; /* n_atoms */
  double coulomb_k;
} ce_molecule_desc;

/* Tolerances: BatchedRK45(mol, k, lanes, rtol, atol) (batched_integrator.hpp:71). */
typedef struct {
  double rtol;
  double atol;
} ce_tolerances;

/* Run configuration: the reference's RunConfig (driver.hpp:28-50) plus the
 * engine's added fields.  ce_run_config_default() fills the reference defaults
 * (pe_stop_fraction 1e-9, redistribute on, max_dt 1e12, max_steps 1e6, Hairer
 * start, 100 rejections, f32). */
typedef struct {
  double pe_stop_fraction;
  double max_dt;
  double t_max;
  double momentum_tol;
  double h0;
  uint64_t max_steps;
  int32_t stop_rule;
  int32_t redistribute_energy;
  int32_t h0_mode;
  int32_t max_rejections;
  int32_t precision;
  int32_t resume;  /* 1: continue every sim from the saved state (ce_engine_enable_state) */
} ce_run_config;

/* Result buffers, caller-allocated host memory; a NULL field is not copied.
 * momenta mirrors BatchedRK45::Result::momenta [(sim*n + i)*3 + c]
 * (batched_integrator.hpp:64-69); momenta_f32 is the same in float (the
 * dataset record's final momenta). */
typedef struct {
  double* momenta;
  float* momenta_f32;
  uint32_t* steps;     /* accepted steps of the sim */
  uint32_t* attempts;  /* attempted steps (accepted + rejected) */
  uint8_t* status;     /* ce_sim_status */
  double* t_final;     /* FP64 time at the stop */
  double* energy_err;  /* (KE + PE - E0)/E0 at the stop, FP64, before redistribution;
                          only after a run with ce_engine_enable_energy_diagnostic on */
  double* trace_h;     /* [sim][trace_cap] accepted step sizes (0 past the end) */
  uint32_t trace_cap;  /* 0 = no trace; must match ce_engine_enable_trace */
  uint32_t pad_;
} ce_results;

/* Totals of the last run, measured on the device. */
typedef struct {
  double kernel_ms;          /* integration kernel, HIP events on the engine stream */
  double sampler_ms;         /* last device-sampler kernel */
  double pack_ms;            /* last record-pack kernel */
  uint64_t sims;
  uint64_t attempted_steps;  /* summed over sims */
  uint64_t accepted_steps;
  uint64_t refills;          /* wave refill events */
  uint64_t lane_slots;       /* simulation seats x integration-loop iterations of
                                every wave (64 seats per wave; 32 where a simulation
                                spans a pair of lanes, f32 N = 8): the share
                                attempted_steps / lane_slots of seats did a step
                                (the rest idled: refill wait, tails) */
  int32_t grid_blocks;
  int32_t block_threads;
  int32_t n_cus;
  int32_t launches;          /* integration kernels launched (1 unless split by size) */
} ce_run_stats;

/* One molecule type of a mixed-size engine (no reference counterpart: the
 * reference runs one fixed molecule per batch, batched_force.hpp:9-10).  Its
 * `count` sims occupy consecutive global indices after the previous species'.
 * charge_levels = 0 uses `charges` for every sim; L > 0 gives each sim its own
 * charges, drawn uniformly from {1..L} by the device sampler (redrawn while
 * their sum exceeds max_total_charge, if that is > 0). */
#define CE_MAX_SPECIES 4
typedef struct {
  int32_t n_atoms;
  int32_t charge_levels;
  const double* masses;   /* n_atoms, electron masses */
  const double* charges;  /* n_atoms (may be NULL when charge_levels > 0) */
  double radius;          /* device-sampler ball radius, bohr */
  double min_separation;
  uint64_t count;
  int32_t max_total_charge;
  int32_t pad_;
} ce_species_desc;

/* Launch shape: 0 fields mean "engine default".  refill_threshold = idle lanes a
 * wave waits for before it pulls new sims (1..64); chunk = sims a wave claims
 * from the device work queue per atomic.  mixed_launch, for a mixed engine of
 * sizes (2, 3, 5, 8): CE_MIXED_ONE_LAUNCH runs all species in one persistent
 * kernel (compiled for the largest size's registers: one wave per SIMD),
 * CE_MIXED_PER_SPECIES one kernel per species on the engine stream (each at its
 * own occupancy, one tail each), CE_MIXED_AUTO (0) the latter when every species
 * has at least 2^19 sims, else the former. */
#define CE_MIXED_AUTO 0
#define CE_MIXED_ONE_LAUNCH 1
#define CE_MIXED_PER_SPECIES 2
typedef struct {
  int32_t block_threads;     /* multiple of 64, <= 256 */
  int32_t blocks_per_cu;     /* resident blocks per CU to launch */
  int32_t refill_threshold;
  int32_t chunk;
  int32_t mixed_launch;      /* CE_MIXED_* */
} ce_launch_config;

typedef struct ce_engine ce_engine;

const char* ce_last_error(void);
int ce_device_count(int* count);
void ce_run_config_default(ce_run_config* cfg);

/* Replaces BatchedRK45(const Molecule&, Real coulomb_k, size_t lanes, T rtol,
 * T atol) (batched_integrator.hpp:71-106).  `capacity` is the most sims one run
 * holds (the reference's `lanes` is the SIMD width; here it is the batch). */
int ce_engine_create(const ce_molecule_desc* mol, const ce_tolerances* tol, size_t capacity,
                     ce_engine** out);
int ce_engine_destroy(ce_engine* eng);
/* A mixed-size engine (1..CE_MAX_SPECIES species) run in one launch: the
 * species occupy global sims species-major; geometry and momenta arrays are
 * species-major with each species' own atom count.  Sizes (2, 3, 5, 8) in that
 * order run as one persistent kernel; other combinations run one kernel per
 * species on the engine stream. */
int ce_engine_create_mixed(const ce_species_desc* species, int32_t n_species, double coulomb_k,
                           const ce_tolerances* tol, ce_engine** out);
/* Device sampler for every species of a mixed engine: geometry from each
 * species' radius / min_separation, and per-sim charges where charge_levels > 0;
 * sim g of the engine has global index first_index + g. */
int ce_engine_sample_species(ce_engine* eng, uint64_t seed, uint64_t first_index,
                             uint32_t max_attempts);
/* Per-sim charges for the loaded sims ([sim][atom], species-major like the
 * geometry), replacing the molecule's fixed charges on the next runs. */
int ce_engine_set_charges(ce_engine* eng, const double* charges, size_t n_sims);
int ce_engine_get_charges(ce_engine* eng, double* charges, size_t n_sims);

/* Replaces set_lane_geometry(lane, positions) (batched_integrator.hpp:112-121)
 * for sims [0, n_sims): pos is AoS doubles [(sim*n + i)*3 + c], velocities start
 * at rest.  The f32 path narrows to float on the device as the reference does. */
int ce_engine_set_geometries(ce_engine* eng, const double* pos, size_t n_sims);
/* Device-side sampler (no reference counterpart; replaces the host loop of
 * UniformSphereSampler::sample, sampler.cpp:28-57, by a counter-based one):
 * sims [0, n_sims) get global indices first_index + s, so results do not depend
 * on how a job is split.  Sims that cannot be placed get CE_SIM_SAMPLER_FAIL. */
int ce_engine_sample(ce_engine* eng, uint64_t seed, uint64_t first_index, size_t n_sims,
                     double radius, double min_sep, uint32_t max_attempts);
/* Copies the current geometries back (AoS doubles). */
int ce_engine_get_geometries(ce_engine* eng, double* pos, size_t n_sims);

/* Replaces BatchedRK45::run(const RunConfig&) (batched_integrator.hpp:123-191):
 * integrates every loaded sim to its stop and copies the requested results
 * into `out` (may be NULL: results stay on the device). */
int ce_engine_run(ce_engine* eng, const ce_run_config* cfg, ce_results* out);
/* One batch as a HIP graph (no reference counterpart; the reference generator
 * samples and integrates batch by batch, examples/generate_dataset.cpp:180-217):
 * ce_engine_capture records device sampling of n_sims sims (as ce_engine_sample),
 * the counter reset and the integration with cfg on the engine's stream, with
 * no memset or copy nodes; ce_engine_replay launches it for sims
 * [first_index, first_index + n_sims) and copies results like ce_engine_run.
 * Single-species engines; a new capture replaces the old one. */
int ce_engine_capture(ce_engine* eng, const ce_run_config* cfg, uint64_t seed, size_t n_sims,
                      double radius, double min_sep, uint32_t max_attempts);
int ce_engine_replay(ce_engine* eng, uint64_t first_index, ce_results* out);
/* Copies the last run's results (same buffers as ce_engine_run). */
int ce_engine_download(ce_engine* eng, ce_results* out);
int ce_engine_stats(const ce_engine* eng, ce_run_stats* out);
int ce_engine_set_launch(ce_engine* eng, const ce_launch_config* lc);
int ce_engine_get_launch(const ce_engine* eng, ce_launch_config* lc);
/* Records the first trace_cap accepted step sizes of every sim on the next
 * runs (the reference's StepObserver hook, driver.hpp:12-18); 0 disables. */
int ce_engine_enable_trace(ce_engine* eng, uint32_t trace_cap);
/* Energy diagnostic (no reference counterpart: BatchedRK45::Result has no energy
 * error): with on != 0 the next runs compute (KE + PE - E0)/E0 per sim in FP64
 * at the stop, for ce_results.energy_err.  Off by default: it adds an FP64
 * potential-energy sum to every sim's finalize. */
int ce_engine_enable_energy_diagnostic(ce_engine* eng, int32_t on);

/* Saved integration state (no reference counterpart; the learned closure stops
 * sims at an early t_max and continues some of them).  When enabled, every run
 * keeps each sim's positions and velocities at its stop (in the run's
 * precision), the next step size and E0 on the device; a later run with
 * cfg->resume = 1 continues each sim from there (time, step counts and the
 * energy reference carry over), e.g. from a t_cut stop to the full stop rule. */
int ce_engine_enable_state(ce_engine* eng, int32_t enable);
/* Host copy of the saved state: pos/vel [sim][n][3], h and e0 per sim (any may
 * be NULL); n_sims must equal the loaded sims. */
int ce_engine_get_state(ce_engine* eng, double* pos, double* vel, double* h, double* e0,
                        size_t n_sims);
/* Sets the state the next run with cfg->resume = 1 starts from, for sims
 * [0, n_sims) (a mixed engine: all its sims): positions and velocities
 * [sim][n][3] in `precision` (CE_PRECISION_*; the run must use the same), per-sim
 * step size h (NULL or <= 0: the run picks the first step as for a fresh sim:
 * fixed h0 or the Hairer heuristic) and E0 (the energy the PE rule's threshold
 * and the redistribution refer to; NULL = 0); step counts and time restart at 0.
 * Enables the saved state.  No reference counterpart as a C entry point: the C++
 * Integrator::step (integrator.hpp:28) and run_to_convergence (driver.hpp:68) of
 * include/coulomb_gpu.hpp advance arbitrary states through it. */
int ce_engine_set_state(ce_engine* eng, const double* pos, const double* vel, const double* h,
                        const double* e0, size_t n_sims, int32_t precision);
/* As ce_engine_run, for the listed sims only (single-species engines); the other
 * sims' results are left as they were. */
int ce_engine_run_subset(ce_engine* eng, const ce_run_config* cfg, const uint64_t* sims,
                         size_t n, ce_results* out);

/* Results of ce_run_device_noalloc: DEVICE memory the caller allocated, laid
 * out as in ce_results (per-atom arrays [(sim*n + i)*3 + c], per-sim arrays of
 * n_sims entries).  Every field is required. */
typedef struct {
  double* momenta;
  float* momenta_f32;
  uint32_t* steps;
  uint32_t* attempts;
  uint8_t* status;
  double* t_final;
} ce_device_results;

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the integration of ce_engine_create + ce_engine_set_geometries
 * + ce_engine_run (same kernels, same bits per sim) over caller-owned device
 * buffers, queued on `stream` (a hipStream_t; NULL = the null stream) and not
 * waited for.  geom is DEVICE memory, AoS doubles [(sim*n + i)*3 + c]; fixed
 * charges from `mol`; no trace, energy diagnostic or saved state; cfg->resume
 * must be 0.  It calls no hipMalloc / hipFree: the kernels' only scratch memory
 * (the work-queue head and run counters, 64 bytes, whatever n_sims is) is a
 * static device array that every GPU holds until the process ends, looked up
 * once per device on the first call there.  Calls that overlap on one GPU share
 * that array and race: serialize them (one stream, or wait between calls). */
int ce_run_device_noalloc(const ce_molecule_desc* mol, const ce_tolerances* tol,
                          const ce_run_config* cfg, const double* geom, size_t n_sims,
                          const ce_device_results* out, void* stream);

/* Dataset records of the reference generator (examples/generate_dataset.cpp:
 * 198-217; numpy dtype examples/to_parquet.py:34-42): per sim, packed
 * little-endian, no padding -- n*3 f32 initial positions, n*3 f32 final momenta
 * (post-redistribution m*v), u32 accepted steps, u8 converged.  24n + 5 bytes. */
size_t ce_record_bytes(int32_t n_atoms);
/* Packs sims [first, first + n_sims) of the last run into records on the device
 * and copies them to host_out (n_sims * ce_record_bytes(n) bytes). */
int ce_engine_pack_records(ce_engine* eng, uint64_t first, size_t n_sims, uint8_t* host_out);
/* The same, queued: the pack kernel runs on the engine's stream and the copy to
 * host_out on a copy stream of its own, so it overlaps the engine's next run
 * (the reference generator writes each batch's records before sampling the
 * next, examples/generate_dataset.cpp:190-217).  Two device record buffers are
 * used in turn: a third call waits on the device for the first one's copy.
 * host_out should be pinned host memory (ce_host_alloc) and must not be read,
 * written or freed before ce_engine_records_wait returns. */
int ce_engine_pack_records_async(ce_engine* eng, uint64_t first, size_t n_sims,
                                 uint8_t* host_out);
/* Waits until every queued record copy has reached host memory. */
int ce_engine_records_wait(ce_engine* eng);
/* Pinned (page-locked) host memory, for buffers the device copies into
 * asynchronously and at the link's full rate; ce_host_free releases it. */
int ce_host_alloc(size_t bytes, void** out);
int ce_host_free(void* p);

/* ---- closure network (no reference counterpart) --------------------------
 * A multilayer perceptron: widths[0] inputs, SiLU hidden layers, widths[n-1]
 * linear outputs; parameters W_0 [w1][w0], b_0 [w1], W_1, b_1, ... in one FP32
 * vector (Glorot-uniform weights from `seed`, zero biases).  Forward and backward
 * run as BF16 MFMA GEMMs with FP32 accumulation; Adam (beta 0.9 / 0.999,
 * eps 1e-8) updates FP32 master parameters.  Loss: mean squared error. */
typedef struct ce_mlp ce_mlp;
int ce_mlp_create(const int32_t* widths, int32_t n_widths, uint64_t seed, ce_mlp** out);
int ce_mlp_destroy(ce_mlp* m);
int64_t ce_mlp_param_count(const ce_mlp* m);
int ce_mlp_get_params(ce_mlp* m, float* params);
int ce_mlp_set_params(ce_mlp* m, const float* params);  /* also resets Adam */
/* y [batch][w_last] = net(x [batch][w0]); host buffers. */
int ce_mlp_forward(ce_mlp* m, const float* x, size_t batch, float* y);
/* Inference (ce_mlp_forward and the closure) runs networks shaped
 * [w0, H, ..., H, w_out <= 32] with H = 256 (w0 <= 64) or 512 (w0 <= 128) in one fused kernel
 * (activations kept on chip; same bits as the layer-by-layer GEMMs) unless
 * disabled here.  Returns 1 if the fused kernel is in use, 0 if not, < 0 on error. */
int ce_mlp_set_fused(ce_mlp* m, int32_t enable);
/* Loss and d loss / d params for one batch (host buffers). */
int ce_mlp_gradients(ce_mlp* m, const float* x, const float* t, size_t batch, float* grad,
                     float* loss);
/* One Adam step with the gradients of the last ce_mlp_gradients call. */
int ce_mlp_adam_step(ce_mlp* m, float lr);
/* Device-resident training set and one shuffled epoch of Adam steps over it. */
int ce_mlp_set_dataset(ce_mlp* m, const float* x, const float* t, size_t n);
int ce_mlp_train_epoch(ce_mlp* m, size_t batch, float lr, uint64_t shuffle_seed,
                       float* mean_loss);

/* ---- learned asymptotic closure (no reference counterpart) ----------------
 * For one molecule size (single-species engines, f32, momentum-or-t_max rule):
 * integrate to t_cut keeping the state, predict the asymptotic momenta of the
 * sims stopped by the cut as a closed-form straight-line-impulse prior (KE
 * rescaled to E0) plus an MLP ensemble's residual, accept the ensemble mean
 * where the members' per-atom spread is <= gate_tol relative to |p_i|
 * (status CE_SIM_CLOSURE), and continue every other cut sim from its saved
 * state to the stop rule.  Network input: 8n features per sim (per atom its
 * position relative to the centre of mass and m v, then the n charges and n
 * masses), shifted and scaled; output: 3n residual components divided by
 * out_scale. */
typedef struct ce_closure ce_closure;
typedef struct {
  uint64_t sims;              /* loaded sims */
  uint64_t candidates;        /* sims stopped by the cut */
  uint64_t fallback;          /* of those, continued by integration */
  uint64_t cut_attempts;      /* attempted steps up to the cut */
  uint64_t fallback_attempts; /* attempted steps of the continuation */
  double cut_ms;              /* integration kernel to t_cut */
  double infer_ms;            /* features + ensemble + gate */
  double fallback_ms;         /* continuation kernel */
  double total_ms;            /* whole ce_closure_run on the device stream */
  uint64_t member_rows;       /* rows the members evaluated, summed over members: with
                                 three or more members the first two decide most sims
                                 (an exact early-rejection bound) and the others run on
                                 the rest only */
} ce_closure_stats;
int ce_closure_create(int32_t n_atoms, const double* masses, const double* charges,
                      double coulomb_k, ce_closure** out);
int ce_closure_destroy(ce_closure* c);
int ce_closure_set_normalization(ce_closure* c, const float* in_shift, const float* in_scale,
                                 const float* out_scale);
/* 2..8 members for a gate (fewer: every cut sim falls back); members are not
 * owned and must outlive their use. */
int ce_closure_set_members(ce_closure* c, ce_mlp* const* members, int32_t n_members);
/* Raw features [sim][8n] and the prior [sim][n][3] of the engine's saved f32
 * state, for every loaded sim (training data). */
int ce_closure_features(ce_closure* c, ce_engine* eng, float* features, double* prior,
                        size_t n_sims);
int ce_closure_run(ce_closure* c, ce_engine* eng, const ce_run_config* cfg, double t_cut,
                   double gate_tol, ce_results* out);
int ce_closure_get_stats(const ce_closure* c, ce_closure_stats* out);

/* Host sampler with the reference's exact stream: UniformSphereSampler(Options,
 * seed) (sampler.hpp:46-72, sampler.cpp:7-57) -- std::mt19937_64 feeding
 * std::uniform_real_distribution<double>(-R, R), cube rejection, sequential
 * min-separation placement.  Stateful like the reference: successive calls
 * continue the stream.  Create fails with CE_ERR_INVALID_ARGUMENT for
 * radius <= 0 or min_separation < 0; sample fails with CE_ERR_RUNTIME when an
 * atom cannot be placed in max_attempts tries (sims before it are written). */
typedef struct ce_host_sampler ce_host_sampler;
int ce_host_sampler_create(uint64_t seed, double radius, double min_separation,
                           uint64_t max_attempts, ce_host_sampler** out);
int ce_host_sampler_reseed(ce_host_sampler* s, uint64_t seed);
int ce_host_sampler_sample(ce_host_sampler* s, int32_t n_atoms, size_t n_sims, double* pos_out);
int ce_host_sampler_destroy(ce_host_sampler* s);

/* Replaces batched::accelerations<float> (batched_force.hpp:31-84) with the same
 * SoA layout p[i*k + lane], pair_const[i*n + j], inv_mass[i]; host pointers. */
int ce_batched_accelerations_f32(size_t n, size_t k, const float* px, const float* py,
                                 const float* pz, const float* pair_const, const float* inv_mass,
                                 float* ax, float* ay, float* az);
/* Replaces batched::potential_energy<float> (batched_force.hpp:89-110). */
int ce_batched_potential_energy_f32(size_t n, size_t k, const float* px, const float* py,
                                    const float* pz, const float* pair_const, float* pe);
/* The same in double: batched::accelerations<double> / potential_energy<double>
 * (batched_force.hpp:31-110); the C++ CoulombForce (system.hpp:25-39) of
 * include/coulomb_gpu.hpp evaluates one configuration through them. */
int ce_batched_accelerations_f64(size_t n, size_t k, const double* px, const double* py,
                                 const double* pz, const double* pair_const,
                                 const double* inv_mass, double* ax, double* ay, double* az);
int ce_batched_potential_energy_f64(size_t n, size_t k, const double* px, const double* py,
                                    const double* pz, const double* pair_const, double* pe);

/* No reference counterpart: a plain device-to-device copy of `bytes` (a
 * multiple of 16) repeated `reps` times on the null stream, average time per
 * copy into *ms_per_copy.  The reference kernel against which the HBM traffic
 * counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) of the engine's kernels are
 * scaled: a copy reads and writes `bytes` each. */
int ce_hbm_copy_calibration(size_t bytes, int32_t reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* COULOMB_EXPLOSION_H */
