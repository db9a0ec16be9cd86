// coulomb_gpu.hpp -- C++ interface of the MI355X engine, header-only over the
// C-ABI in coulomb_explosion.h.
//
// It keeps the reference's strategy API with the same names, argument meaning
// and exception types, so a caller of the reference switches by changing the
// namespace (coulomb:: -> coulomb::gpu::):
//   Real, Vec3, units                     include/coulomb/types.hpp, units.hpp
//   Atom / atom_from_amu / Molecule       include/coulomb/molecule.hpp:14-35
//   State, CoulombForce                   include/coulomb/system.hpp:13-39
//   Sampler, UniformSphereSampler         include/coulomb/sampler.hpp:23-72
//   Integrator, IntegratorKind,
//   IntegratorOptions, make_integrator    include/coulomb/integrator.hpp:18-54
//   StepObserver, RunConfig, RunResult,
//   run_to_convergence                    include/coulomb/driver.hpp:18-71
//   BatchedRK45<T>                        include/coulomb/batched_integrator.hpp:62-191
// Every evaluation runs on the GPU through the library: CoulombForce through the
// reference-layout force / PE kernels, the RK45 integrator and
// run_to_convergence through the engine (a one-simulation engine whose saved
// state carries the integrator's state between calls), BatchedRK45 through a
// batch engine.  Deliberate differences:
//   * BatchedRK45's `lanes` is the batch capacity (any size), not the host SIMD
//     width, and each simulation has its own step size (no shared dt), so
//     Result::batch_steps is the largest per-sim step count;
//   * RunConfig carries the engine's extra fields (stop rule, t_max,
//     momentum_tol, fixed h0, max_rejections) with the reference's defaults;
//     run_to_convergence uses the reference's four fields;
//   * make_integrator(VelocityVerlet) throws std::invalid_argument (the
//     symplectic scheme is not part of the engine);
//   * UniformSphereSampler draws the reference's default (Real = double)
//     stream; with COULOMB_SINGLE_PRECISION the positions are that stream
//     narrowed to float.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

#include "coulomb_explosion.h"

namespace coulomb {
namespace gpu {

#ifdef COULOMB_SINGLE_PRECISION
using Real = float;
#else
using Real = double;
#endif

struct Vec3 {
  Real x{0}, y{0}, z{0};
  constexpr Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  constexpr Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  constexpr Vec3 operator*(Real s) const { return {x * s, y * s, z * s}; }
  constexpr Vec3& operator+=(const Vec3& o) {
    x += o.x, y += o.y, z += o.z;
    return *this;
  }
  constexpr Vec3& operator-=(const Vec3& o) {
    x -= o.x, y -= o.y, z -= o.z;
    return *this;
  }
};
constexpr Vec3 operator*(Real s, const Vec3& v) { return v * s; }
constexpr Real dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Real norm(const Vec3& v) { return std::sqrt(dot(v, v)); }

namespace units {
inline constexpr Real kAmuToElectronMass = 1822.888486209;  // CODATA 2018 m_u / m_e
constexpr Real amu_to_electron_mass(Real mass_amu) { return mass_amu * kAmuToElectronMass; }
}  // namespace units

struct Atom {
  std::string symbol;
  Real mass{0};    // electron masses
  Real charge{0};
};

inline Atom atom_from_amu(std::string symbol, Real mass_amu, Real charge) {
  return Atom{std::move(symbol), units::amu_to_electron_mass(mass_amu), charge};
}

struct Molecule {
  std::vector<Atom> atoms;
  std::size_t size() const { return atoms.size(); }
  bool empty() const { return atoms.empty(); }
};

// Positions and velocities, one entry per atom.
struct State {
  std::vector<Vec3> positions;
  std::vector<Vec3> velocities;
  std::size_t size() const { return positions.size(); }
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
template <class T>
constexpr int32_t precision_of() {
  return std::is_same<T, float>::value ? CE_PRECISION_F32 : CE_PRECISION_F64;
}
inline void flatten(const std::vector<Vec3>& v, double* out) {
  for (std::size_t i = 0; i < v.size(); ++i) {
    out[3 * i] = static_cast<double>(v[i].x);
    out[3 * i + 1] = static_cast<double>(v[i].y);
    out[3 * i + 2] = static_cast<double>(v[i].z);
  }
}
inline void unflatten(const double* in, std::size_t n, std::vector<Vec3>& v) {
  v.resize(n);
  for (std::size_t i = 0; i < n; ++i)
    v[i] = Vec3{static_cast<Real>(in[3 * i]), static_cast<Real>(in[3 * i + 1]),
                static_cast<Real>(in[3 * i + 2])};
}
// The reference-layout kernels by element type.
inline int soa_accelerations(std::size_t n, const float* px, const float* py, const float* pz,
                             const float* pc, const float* im, float* ax, float* ay, float* az) {
  return ce_batched_accelerations_f32(n, 1, px, py, pz, pc, im, ax, ay, az);
}
inline int soa_accelerations(std::size_t n, const double* px, const double* py, const double* pz,
                             const double* pc, const double* im, double* ax, double* ay,
                             double* az) {
  return ce_batched_accelerations_f64(n, 1, px, py, pz, pc, im, ax, ay, az);
}
inline int soa_potential_energy(std::size_t n, const float* px, const float* py, const float* pz,
                                const float* pc, float* pe) {
  return ce_batched_potential_energy_f32(n, 1, px, py, pz, pc, pe);
}
inline int soa_potential_energy(std::size_t n, const double* px, const double* py,
                                const double* pz, const double* pc, double* pe) {
  return ce_batched_potential_energy_f64(n, 1, px, py, pz, pc, pe);
}
inline void masses_charges(const Molecule& mol, std::vector<double>& m, std::vector<double>& q) {
  m.resize(mol.size());
  q.resize(mol.size());
  for (std::size_t i = 0; i < mol.size(); ++i) {
    m[i] = static_cast<double>(mol.atoms[i].mass);
    q[i] = static_cast<double>(mol.atoms[i].charge);
  }
}
// A simulation engine of the library with its device buffers (move-only).
class EngineHandle {
 public:
  EngineHandle() = default;
  EngineHandle(const Molecule& mol, double coulomb_k, std::size_t capacity, double rtol,
               double atol) {
    masses_charges(mol, m_, q_);
    ce_molecule_desc d{static_cast<int32_t>(m_.size()), 0, m_.data(), q_.data(), coulomb_k};
    ce_tolerances t{rtol, atol};
    check(ce_engine_create(&d, &t, capacity, &h_));
    k_ = coulomb_k;
    rtol_ = rtol;
    atol_ = atol;
  }
  ~EngineHandle() { ce_engine_destroy(h_); }
  EngineHandle(const EngineHandle&) = delete;
  EngineHandle& operator=(const EngineHandle&) = delete;
  EngineHandle(EngineHandle&& o) noexcept { *this = std::move(o); }
  EngineHandle& operator=(EngineHandle&& o) noexcept {
    std::swap(h_, o.h_);
    std::swap(m_, o.m_);
    std::swap(q_, o.q_);
    std::swap(k_, o.k_);
    std::swap(rtol_, o.rtol_);
    std::swap(atol_, o.atol_);
    return *this;
  }
  ce_engine* get() const { return h_; }
  // Whether this engine was made for `mol` with these constants.
  bool matches(const Molecule& mol, double coulomb_k, double rtol, double atol) const {
    if (h_ == nullptr || mol.size() != m_.size() || coulomb_k != k_ || rtol != rtol_ ||
        atol != atol_)
      return false;
    for (std::size_t i = 0; i < m_.size(); ++i)
      if (static_cast<double>(mol.atoms[i].mass) != m_[i] ||
          static_cast<double>(mol.atoms[i].charge) != q_[i])
        return false;
    return true;
  }

 private:
  ce_engine* h_{nullptr};
  std::vector<double> m_, q_;
  double k_{0}, rtol_{0}, atol_{0};
};
}  // namespace detail

// Coulomb interaction between point charges (system.hpp:25-39), evaluated on
// the GPU by the reference-layout kernels (one configuration, one lane).
class CoulombForce {
 public:
  explicit CoulombForce(Real coulomb_constant = Real{1}) : k_(coulomb_constant) {}

  // Accelerations a_i = sum_j k q_i q_j (x_i - x_j) / |x_i - x_j|^3 / m_i into
  // `forces` (resized to N), as the reference's method of this name returns.
  void accelerations(const Molecule& molecule, const State& state,
                     std::vector<Vec3>& forces) const {
    const std::size_t n = molecule.size();
    forces.assign(n, Vec3{});
    if (n == 0) return;
    std::vector<Real> p[3], a[3], pc(n * n), im(n);
    for (int c = 0; c < 3; ++c) {
      p[c].resize(n);
      a[c].resize(n);
    }
    fill(molecule, state, p, pc.data());
    for (std::size_t i = 0; i < n; ++i) im[i] = Real{1} / molecule.atoms[i].mass;
    detail::check(detail::soa_accelerations(n, p[0].data(), p[1].data(), p[2].data(), pc.data(),
                                            im.data(), a[0].data(), a[1].data(), a[2].data()));
    for (std::size_t i = 0; i < n; ++i) forces[i] = Vec3{a[0][i], a[1][i], a[2][i]};
  }

  Real potential_energy(const Molecule& molecule, const State& state) const {
    const std::size_t n = molecule.size();
    if (n < 2) return Real{0};
    std::vector<Real> p[3], pc(n * n);
    for (int c = 0; c < 3; ++c) p[c].resize(n);
    fill(molecule, state, p, pc.data());
    Real pe{0};
    detail::check(detail::soa_potential_energy(n, p[0].data(), p[1].data(), p[2].data(),
                                               pc.data(), &pe));
    return pe;
  }

  // sum_i m_i |v_i|^2 / 2 (bookkeeping of N terms, on the host).
  static Real kinetic_energy(const Molecule& molecule, const State& state) {
    Real ke{0};
    for (std::size_t i = 0; i < molecule.size(); ++i)
      ke += Real{0.5} * molecule.atoms[i].mass * dot(state.velocities[i], state.velocities[i]);
    return ke;
  }

  Real coulomb_constant() const { return k_; }

 private:
  void fill(const Molecule& mol, const State& st, std::vector<Real>* p, Real* pc) const {
    const std::size_t n = mol.size();
    if (st.positions.size() != n)
      throw std::invalid_argument("CoulombForce: state and molecule sizes differ");
    for (std::size_t i = 0; i < n; ++i) {
      p[0][i] = st.positions[i].x;
      p[1][i] = st.positions[i].y;
      p[2][i] = st.positions[i].z;
      for (std::size_t j = 0; j < n; ++j)
        pc[i * n + j] = static_cast<Real>(static_cast<double>(k_) *
                                          static_cast<double>(mol.atoms[i].charge) *
                                          static_cast<double>(mol.atoms[j].charge));
    }
  }

  Real k_;
};

// Strategy interface for initial-geometry samplers (sampler.hpp:23-34).
class Sampler {
 public:
  virtual ~Sampler() = default;
  // One configuration for `molecule` into `state`: positions resized and
  // filled, velocities resized and zeroed.
  virtual void sample(const Molecule& molecule, State& state) = 0;
  virtual std::string_view name() const = 0;
};

// Positions uniform in a ball, pairs closer than min_separation redrawn by
// sequential placement (sampler.hpp:46-72): the reference's std::mt19937_64
// stream for a seed, drawn by the library's host code.
class UniformSphereSampler : public Sampler {
 public:
  struct Options {
    Real radius{4};
    Real min_separation{0.25};
    std::size_t max_attempts_per_atom{10000};
  };

  UniformSphereSampler(Options options, std::uint64_t seed) : opts_(options) {
    detail::check(ce_host_sampler_create(seed, opts_.radius, opts_.min_separation,
                                         opts_.max_attempts_per_atom, &h_));
  }
  ~UniformSphereSampler() override { ce_host_sampler_destroy(h_); }
  UniformSphereSampler(const UniformSphereSampler&) = delete;
  UniformSphereSampler& operator=(const UniformSphereSampler&) = delete;

  void sample(const Molecule& molecule, State& state) override {
    std::vector<double> pos(3 * molecule.size());
    detail::check(ce_host_sampler_sample(h_, static_cast<int32_t>(molecule.size()), 1,
                                         pos.data()));
    detail::unflatten(pos.data(), molecule.size(), state.positions);
    state.velocities.assign(molecule.size(), Vec3{});
  }
  std::string_view name() const override { return "uniform-sphere"; }
  void reseed(std::uint64_t seed) { detail::check(ce_host_sampler_reseed(h_, seed)); }
  const Options& options() const { return opts_; }

 private:
  Options opts_;
  ce_host_sampler* h_{nullptr};
};

// Strategy interface for time integrators (integrator.hpp:18-33).
class Integrator {
 public:
  virtual ~Integrator() = default;
  // Advances `state` by one step no larger than dt; returns the step taken.
  virtual Real step(const Molecule& molecule, const CoulombForce& force, State& state,
                    Real dt) = 0;
  virtual std::string_view name() const = 0;
};

enum class IntegratorKind { VelocityVerlet, RK45 };

struct IntegratorOptions {
  Real rtol{1e-8};
  Real atol{1e-16};
};

// Adaptive DP5(4) of the reference (integrators.cpp:59-250) on the GPU engine.
// Like the reference it keeps its step size between calls (the first call picks
// it with the Hairer heuristic); each call uploads the state to a
// one-simulation engine and integrates exactly one accepted step there.
class DormandPrince45 final : public Integrator {
 public:
  explicit DormandPrince45(const IntegratorOptions& o) : opts_(o) {}

  Real step(const Molecule& molecule, const CoulombForce& force, State& state,
            Real dt) override {
    const std::size_t n = molecule.size();
    if (state.positions.size() != n || state.velocities.size() != n)
      throw std::invalid_argument("DormandPrince45: state and molecule sizes differ");
    ce_engine* e = engine(molecule, force);
    std::vector<double> pos(3 * n), vel(3 * n);
    detail::flatten(state.positions, pos.data());
    detail::flatten(state.velocities, vel.data());
    double h = static_cast<double>(h_);
    detail::check(ce_engine_set_state(e, pos.data(), vel.data(), &h, nullptr, 1, kPrecision));
    ce_run_config c;
    ce_run_config_default(&c);
    c.stop_rule = CE_STOP_MOMENTUM_OR_TMAX;  // ends only at max_steps here:
    c.momentum_tol = 0.0;                    // no momentum stop,
    c.t_max = 1e300;                         // no time limit,
    c.max_steps = 1;                         // one accepted step
    c.max_dt = static_cast<double>(dt);
    c.redistribute_energy = 0;
    c.h0_mode = CE_H0_HAIRER;
    c.resume = 1;
    c.precision = kPrecision;
    uint8_t status = 0;
    double taken = 0.0;
    ce_results out{};
    out.status = &status;
    out.t_final = &taken;
    detail::check(ce_engine_run(e, &c, &out));
    if (status == CE_SIM_REJECT_LIMIT || status == CE_SIM_NONFINITE)
      throw std::runtime_error(
          "Dormand-Prince: step rejected too many times (singular collision or tolerance too "
          "tight?)");
    detail::check(ce_engine_get_state(e, pos.data(), vel.data(), &h, nullptr, 1));
    detail::unflatten(pos.data(), n, state.positions);
    detail::unflatten(vel.data(), n, state.velocities);
    h_ = static_cast<Real>(h);
    return static_cast<Real>(taken);
  }

  std::string_view name() const override { return "dormand-prince-45"; }

  const IntegratorOptions& options() const { return opts_; }
  // The step size the next step starts from (0 before the first step).
  Real step_size() const { return h_; }
  void set_step_size(Real h) { h_ = h; }
  // The one-simulation engine for `molecule` (made or remade on demand).
  ce_engine* engine(const Molecule& molecule, const CoulombForce& force) {
    const double k = static_cast<double>(force.coulomb_constant());
    if (!eng_.matches(molecule, k, opts_.rtol, opts_.atol)) {
      eng_ = detail::EngineHandle(molecule, k, 1, opts_.rtol, opts_.atol);
      h_ = Real{0};
    }
    return eng_.get();
  }
  static constexpr int32_t kPrecision = detail::precision_of<Real>();

 private:
  IntegratorOptions opts_;
  Real h_{0};
  detail::EngineHandle eng_;
};

inline std::unique_ptr<Integrator> make_integrator(IntegratorKind kind,
                                                   const IntegratorOptions& options) {
  switch (kind) {
    case IntegratorKind::RK45:
      return std::make_unique<DormandPrince45>(options);
    case IntegratorKind::VelocityVerlet:
      throw std::invalid_argument("make_integrator: VelocityVerlet is not provided by the GPU engine");
  }
  throw std::invalid_argument("make_integrator: unknown integrator kind");
}
inline std::unique_ptr<Integrator> make_integrator(IntegratorKind kind) {
  return make_integrator(kind, IntegratorOptions{});
}

// Per-step hook: (0-based step index, time at the start of the step, step taken).
using StepObserver = std::function<void(std::size_t step_index, Real t_start, Real dt)>;

enum class StopRule { PeFraction = CE_STOP_PE_FRACTION, MomentumOrTmax = CE_STOP_MOMENTUM_OR_TMAX };
enum class InitialStep { Hairer = CE_H0_HAIRER, Fixed = CE_H0_FIXED };

// driver.hpp:28-50 (the first four fields), plus the engine's fields, used by
// BatchedRK45.
struct RunConfig {
  Real pe_stop_fraction{1e-9};
  bool redistribute_energy{true};
  Real max_dt{1e12};
  std::size_t max_steps{1'000'000};
  StopRule stop_rule{StopRule::PeFraction};
  double t_max{1e4};
  double momentum_tol{1e-7};
  InitialStep initial_step{InitialStep::Hairer};
  double h0{1e-3};
  int max_rejections{100};
};

// driver.hpp:53-62.
struct RunResult {
  std::size_t steps{0};
  Real t_final{0};
  Real energy_initial{0};
  Real pe_final{0};
  Real ke_before_redist{0};
  Real ke_after_redist{0};
  Real redistribution_scale{1};
  bool converged{false};
};

namespace detail {
// The redistribution at the end of a run (driver.hpp:36-42).
inline void finish_run(const Molecule& mol, const RunConfig& cfg, State& state,
                       RunResult& r) {
  r.ke_before_redist = CoulombForce::kinetic_energy(mol, state);
  if (cfg.redistribute_energy && r.ke_before_redist > Real{0}) {
    const Real s = std::sqrt(r.energy_initial / r.ke_before_redist);
    for (auto& v : state.velocities) v = v * s;
    r.redistribution_scale = s;
  }
  r.ke_after_redist = CoulombForce::kinetic_energy(mol, state);
}
}  // namespace detail

// Integrates `state` until PE <= pe_stop_fraction * PE0 (or max_steps), then
// redistributes the remaining potential energy (driver.hpp:64-71): the
// integrator's steps through `integrator`, the energies through `force`.  A
// state at rest given to a fresh GPU DormandPrince45 runs as one engine launch
// instead (same method, same stop and result; the observer is fed from the
// engine's step trace).
inline RunResult run_to_convergence(const Molecule& molecule, const CoulombForce& force,
                                    Integrator& integrator, State& state,
                                    const RunConfig& config, const StepObserver& observer = {}) {
  RunResult r;
  auto* gpu = dynamic_cast<DormandPrince45*>(&integrator);
  bool at_rest = true;
  for (const Vec3& v : state.velocities) at_rest = at_rest && dot(v, v) == Real{0};
  constexpr uint32_t kTraceCap = 4096;
  const std::size_t n = molecule.size();
  if (gpu != nullptr && at_rest && gpu->step_size() == Real{0} && n >= 1 &&
      (!observer || config.max_steps <= kTraceCap)) {
    ce_engine* e = gpu->engine(molecule, force);
    std::vector<double> pos(3 * n), vel(3 * n);
    detail::flatten(state.positions, pos.data());
    detail::check(ce_engine_set_geometries(e, pos.data(), 1));
    detail::check(ce_engine_enable_state(e, 1));
    detail::check(ce_engine_enable_trace(e, observer ? kTraceCap : 0));
    ce_run_config c;
    ce_run_config_default(&c);
    c.stop_rule = CE_STOP_PE_FRACTION;
    c.pe_stop_fraction = static_cast<double>(config.pe_stop_fraction);
    c.max_dt = static_cast<double>(config.max_dt);
    c.max_steps = config.max_steps;
    c.redistribute_energy = 0;  // done below, with the reference's bookkeeping
    c.precision = DormandPrince45::kPrecision;
    uint32_t steps = 0;
    uint8_t status = 0;
    double t_final = 0.0;
    std::vector<double> trace(observer ? kTraceCap : 0);
    ce_results out{};
    out.steps = &steps;
    out.status = &status;
    out.t_final = &t_final;
    out.trace_h = observer ? trace.data() : nullptr;
    out.trace_cap = observer ? kTraceCap : 0;
    detail::check(ce_engine_run(e, &c, &out));
    if (status == CE_SIM_REJECT_LIMIT || status == CE_SIM_NONFINITE)
      throw std::runtime_error("Dormand-Prince: step rejected too many times");
    double h = 0.0, e0 = 0.0;
    detail::check(ce_engine_get_state(e, pos.data(), vel.data(), &h, &e0, 1));
    detail::unflatten(pos.data(), n, state.positions);
    detail::unflatten(vel.data(), n, state.velocities);
    gpu->set_step_size(static_cast<Real>(h));
    r.steps = steps;
    r.energy_initial = static_cast<Real>(e0);
    // the reference leaves `converged` false when PE0 is already below the
    // threshold (no step taken)
    r.converged = status == CE_SIM_OK_PE && steps > 0;
    if (observer) {
      Real t{0};
      for (std::size_t i = 0; i < steps; ++i) {
        const Real dt = static_cast<Real>(trace[i]);
        observer(i, t, dt);
        t += dt;
      }
      r.t_final = t;
    } else {
      r.t_final = static_cast<Real>(t_final);
    }
    r.pe_final = force.potential_energy(molecule, state);
    detail::finish_run(molecule, config, state, r);
    return r;
  }
  // Any integrator, any state: the reference driver's loop.
  const Real pe0 = force.potential_energy(molecule, state);
  r.energy_initial = pe0 + CoulombForce::kinetic_energy(molecule, state);
  const Real threshold = config.pe_stop_fraction * pe0;
  Real pe = pe0;
  while (pe > threshold) {
    if (r.steps >= config.max_steps) {
      r.converged = false;
      break;
    }
    const Real dt = integrator.step(molecule, force, state, config.max_dt);
    if (observer) observer(r.steps, r.t_final, dt);
    r.t_final += dt;
    ++r.steps;
    pe = force.potential_energy(molecule, state);
    r.converged = true;
  }
  r.pe_final = pe;
  detail::finish_run(molecule, config, state, r);
  return r;
}

// Drop-in for BatchedRK45<T> (batched_integrator.hpp:62-191): T = float is the
// engine's f32 path, T = double its f64 path.
template <class T>
class BatchedRK45 {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "BatchedRK45<T>: T is float or double");

 public:
  struct Result {
    std::vector<double> momenta;        // [(lane*n + i)*3 + c] = m_i s v_i
    std::vector<std::size_t> steps;     // accepted steps per sim
    std::vector<char> converged;        // 1 unless not converged / failed
    std::size_t batch_steps{0};         // max steps over sims (no shared dt)
    std::vector<std::uint8_t> status;   // ce_sim_status per sim
    std::vector<double> t_final;
  };

  BatchedRK45(const Molecule& mol, Real coulomb_k, std::size_t lanes, T rtol, T atol)
      : n_(mol.size()), k_(lanes), geom_(lanes * mol.size() * 3, 0.0) {
    std::vector<double> m, q;
    detail::masses_charges(mol, m, q);
    ce_molecule_desc d{static_cast<int32_t>(n_), 0, m.data(), q.data(),
                       static_cast<double>(coulomb_k)};
    ce_tolerances t{static_cast<double>(rtol), static_cast<double>(atol)};
    detail::check(ce_engine_create(&d, &t, k_, &h_));
  }
  ~BatchedRK45() { ce_engine_destroy(h_); }
  BatchedRK45(const BatchedRK45&) = delete;
  BatchedRK45& operator=(const BatchedRK45&) = delete;

  std::size_t lanes() const { return k_; }
  std::size_t atoms() const { return n_; }

  // Lane `lane`'s initial positions (velocities start at rest).
  void set_lane_geometry(std::size_t lane, const std::vector<Vec3>& pos) {
    if (lane >= k_ || pos.size() != n_) throw std::invalid_argument("set_lane_geometry");
    detail::flatten(pos, geom_.data() + lane * n_ * 3);
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
    c.pe_stop_fraction = static_cast<double>(cfg.pe_stop_fraction);
    c.redistribute_energy = cfg.redistribute_energy ? 1 : 0;
    c.max_dt = static_cast<double>(cfg.max_dt);
    c.max_steps = cfg.max_steps;
    c.stop_rule = static_cast<int32_t>(cfg.stop_rule);
    c.t_max = cfg.t_max;
    c.momentum_tol = cfg.momentum_tol;
    c.h0_mode = static_cast<int32_t>(cfg.initial_step);
    c.h0 = cfg.h0;
    c.max_rejections = cfg.max_rejections;
    c.precision = detail::precision_of<T>();
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
