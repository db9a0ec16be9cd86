// coulomb_gpu.hpp -- C++ interface of the MI355X engine, header-only over the
// C-ABI in coulomb_explosion.h.
//
// It keeps the reference's strategy API for the batched f32 path, with the same
// names, argument meaning and exception types:
//   coulomb::gpu::Atom / atom_from_amu / Molecule     include/coulomb/molecule.hpp:14-35
//   coulomb::gpu::RunConfig                           include/coulomb/driver.hpp:28-50
//   coulomb::gpu::UniformSphereSampler                include/coulomb/sampler.hpp:46-72
//   coulomb::gpu::BatchedRK45                         include/coulomb/batched_integrator.hpp:62-191
// so a caller of BatchedRK45<float> (e.g. examples/generate_dataset.cpp:168-227)
// switches by changing the namespace.  Deliberate differences:
//   * `lanes` is the batch capacity (any size), not the host SIMD width;
//   * each simulation has its own step size (no shared dt), so
//     Result::batch_steps is the largest per-sim step count;
//   * RunConfig carries the engine's extra fields (stop rule, t_max,
//     momentum_tol, fixed h0, precision) with the reference's defaults.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "coulomb_explosion.h"

namespace coulomb {
namespace gpu {

struct Vec3 {
  double x{0}, y{0}, z{0};
};

inline constexpr double kAmuToElectronMass = 1822.888486209;  // CODATA 2018 m_u / m_e

struct Atom {
  std::string symbol;
  double mass{0};    // electron masses
  double charge{0};
};

inline Atom atom_from_amu(std::string symbol, double mass_amu, double charge) {
  return Atom{std::move(symbol), mass_amu * kAmuToElectronMass, charge};
}

struct Molecule {
  std::vector<Atom> atoms;
  std::size_t size() const { return atoms.size(); }
  bool empty() const { return atoms.empty(); }
};

enum class StopRule { PeFraction = CE_STOP_PE_FRACTION, MomentumOrTmax = CE_STOP_MOMENTUM_OR_TMAX };
enum class InitialStep { Hairer = CE_H0_HAIRER, Fixed = CE_H0_FIXED };
enum class Precision { F32 = CE_PRECISION_F32, F64 = CE_PRECISION_F64 };

struct RunConfig {
  double pe_stop_fraction{1e-9};
  bool redistribute_energy{true};
  double max_dt{1e12};
  std::size_t max_steps{1'000'000};
  StopRule stop_rule{StopRule::PeFraction};
  double t_max{1e4};
  double momentum_tol{1e-7};
  InitialStep initial_step{InitialStep::Hairer};
  double h0{1e-3};
  int max_rejections{100};
  Precision precision{Precision::F32};
};

namespace detail {
[[noreturn]] inline void raise(int rc) {
  const std::string msg = ce_last_error();
  if (rc == CE_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}
inline void check(int rc) {
  if (rc != CE_OK) raise(rc);
}
}  // namespace detail

class UniformSphereSampler {
 public:
  struct Options {
    double radius{4};
    double min_separation{0.25};
    std::size_t max_attempts_per_atom{10000};
  };

  UniformSphereSampler(Options options, std::uint64_t seed) : opts_(options) {
    detail::check(ce_host_sampler_create(seed, opts_.radius, opts_.min_separation,
                                         opts_.max_attempts_per_atom, &h_));
  }
  ~UniformSphereSampler() { ce_host_sampler_destroy(h_); }
  UniformSphereSampler(const UniformSphereSampler&) = delete;
  UniformSphereSampler& operator=(const UniformSphereSampler&) = delete;

  // One configuration (positions; atoms at rest).
  std::vector<Vec3> sample(const Molecule& molecule) {
    std::vector<Vec3> pos(molecule.size());
    detail::check(ce_host_sampler_sample(h_, static_cast<int32_t>(molecule.size()), 1,
                                         reinterpret_cast<double*>(pos.data())));
    return pos;
  }
  void reseed(std::uint64_t seed) { detail::check(ce_host_sampler_reseed(h_, seed)); }
  const Options& options() const { return opts_; }
  const char* name() const { return "uniform-sphere"; }

 private:
  Options opts_;
  ce_host_sampler* h_{nullptr};
};

class BatchedRK45 {
 public:
  struct Result {
    std::vector<double> momenta;        // [(lane*n + i)*3 + c] = m_i s v_i
    std::vector<std::size_t> steps;     // accepted steps per sim
    std::vector<char> converged;        // 1 unless not converged / failed
    std::size_t batch_steps{0};         // max steps over sims (no shared dt)
    std::vector<std::uint8_t> status;   // ce_sim_status per sim
    std::vector<double> t_final;
  };

  BatchedRK45(const Molecule& mol, double coulomb_k, std::size_t lanes, double rtol, double atol)
      : n_(mol.size()), k_(lanes), geom_(lanes * mol.size() * 3, 0.0) {
    std::vector<double> m(n_), q(n_);
    for (std::size_t i = 0; i < n_; ++i) {
      m[i] = mol.atoms[i].mass;
      q[i] = mol.atoms[i].charge;
    }
    ce_molecule_desc d{static_cast<int32_t>(n_), 0, m.data(), q.data(), coulomb_k};
    ce_tolerances t{rtol, atol};
    detail::check(ce_engine_create(&d, &t, k_, &h_));
  }
  ~BatchedRK45() { ce_engine_destroy(h_); }
  BatchedRK45(const BatchedRK45&) = delete;
  BatchedRK45& operator=(const BatchedRK45&) = delete;

  std::size_t lanes() const { return k_; }
  std::size_t atoms() const { return n_; }

  void set_lane_geometry(std::size_t lane, const std::vector<Vec3>& pos) {
    if (lane >= k_ || pos.size() != n_) throw std::invalid_argument("set_lane_geometry");
    for (std::size_t i = 0; i < n_; ++i) {
      geom_[(lane * n_ + i) * 3 + 0] = pos[i].x;
      geom_[(lane * n_ + i) * 3 + 1] = pos[i].y;
      geom_[(lane * n_ + i) * 3 + 2] = pos[i].z;
    }
    dirty_ = true;
  }

  // Device-side counter-based sampler for sims [0, n): global index first + s.
  void sample(std::uint64_t seed, std::uint64_t first, std::size_t n,
              const UniformSphereSampler::Options& o) {
    const std::size_t attempts = std::min<std::size_t>(o.max_attempts_per_atom, UINT32_MAX);
    detail::check(ce_engine_sample(h_, seed, first, n, o.radius, o.min_separation,
                                   static_cast<uint32_t>(attempts)));
    loaded_ = n;
    dirty_ = false;
  }

  Result run(const RunConfig& cfg) {
    if (dirty_) {
      detail::check(ce_engine_set_geometries(h_, geom_.data(), k_));
      loaded_ = k_;
      dirty_ = false;
    }
    ce_run_config c;
    ce_run_config_default(&c);
    c.pe_stop_fraction = cfg.pe_stop_fraction;
    c.redistribute_energy = cfg.redistribute_energy ? 1 : 0;
    c.max_dt = cfg.max_dt;
    c.max_steps = cfg.max_steps;
    c.stop_rule = static_cast<int32_t>(cfg.stop_rule);
    c.t_max = cfg.t_max;
    c.momentum_tol = cfg.momentum_tol;
    c.h0_mode = static_cast<int32_t>(cfg.initial_step);
    c.h0 = cfg.h0;
    c.max_rejections = cfg.max_rejections;
    c.precision = static_cast<int32_t>(cfg.precision);
    Result r;
    r.momenta.assign(loaded_ * n_ * 3, 0.0);
    std::vector<uint32_t> steps(loaded_);
    r.status.assign(loaded_, 0);
    r.t_final.assign(loaded_, 0.0);
    ce_results out{};
    out.momenta = r.momenta.data();
    out.steps = steps.data();
    out.status = r.status.data();
    out.t_final = r.t_final.data();
    detail::check(ce_engine_run(h_, &c, &out));
    r.steps.assign(steps.begin(), steps.end());
    r.converged.assign(loaded_, 0);
    for (std::size_t s = 0; s < loaded_; ++s) {
      r.converged[s] = r.status[s] <= CE_SIM_OK_TMAX ? 1 : 0;
      r.batch_steps = std::max<std::size_t>(r.batch_steps, r.steps[s]);
      // The reference throws on a rejection storm (batched_integrator.hpp:269-271).
      if (r.status[s] == CE_SIM_REJECT_LIMIT || r.status[s] == CE_SIM_NONFINITE)
        throw std::runtime_error("BatchedRK45: step rejected too many times");
    }
    return r;
  }

  ce_engine* handle() { return h_; }

 private:
  std::size_t n_, k_;
  std::vector<double> geom_;
  std::size_t loaded_{0};
  bool dirty_{false};
  ce_engine* h_{nullptr};
};

}  // namespace gpu
}  // namespace coulomb
