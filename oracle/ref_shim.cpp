// C entry points over the reference's own scalar engine -- TEST INFRASTRUCTURE ONLY.
//
// This file is a caller of the reference API, compiled together with the
// reference sources where they lie (src/system.cpp, src/integrators.cpp,
// src/driver.cpp, src/sampler.cpp; headers include/coulomb/*.hpp) by
// oracle/Makefile into oracle/_ref/.  Built twice: the default (Real = double)
// and -DCOULOMB_SINGLE_PRECISION (Real = float, include/coulomb/types.hpp:16-20).
// It lets the tests check the oracle's restatement against the reference itself.
#include <cmath>
#include <cstdint>
#include <exception>
#include <stdexcept>
#include <vector>

#include "coulomb/driver.hpp"
#include "coulomb/integrator.hpp"
#include "coulomb/molecule.hpp"
#include "coulomb/sampler.hpp"
#include "coulomb/system.hpp"

using namespace coulomb;

namespace {

Molecule mk(int n, const double* m, const double* q) {
  Molecule mol;
  for (int i = 0; i < n; ++i) mol.atoms.push_back(Atom{"X", static_cast<Real>(m[i]), static_cast<Real>(q[i])});
  return mol;
}

State st(int n, const double* pos) {
  State s;
  for (int i = 0; i < n; ++i) {
    s.positions.push_back(Vec3{static_cast<Real>(pos[3 * i]), static_cast<Real>(pos[3 * i + 1]),
                               static_cast<Real>(pos[3 * i + 2])});
    s.velocities.push_back(Vec3{});
  }
  return s;
}

}  // namespace

extern "C" {

int cer_real_bytes(void) { return static_cast<int>(sizeof(Real)); }

void cer_accelerations(int n, const double* m, const double* q, double k, const double* pos,
                       double* acc) {
  const Molecule mol = mk(n, m, q);
  const State s = st(n, pos);
  std::vector<Vec3> a;
  CoulombForce(static_cast<Real>(k)).accelerations(mol, s, a);
  for (int i = 0; i < n; ++i) {
    acc[3 * i] = a[i].x;
    acc[3 * i + 1] = a[i].y;
    acc[3 * i + 2] = a[i].z;
  }
}

double cer_potential_energy(int n, const double* m, const double* q, double k, const double* pos) {
  return CoulombForce(static_cast<Real>(k)).potential_energy(mk(n, m, q), st(n, pos));
}

// UniformSphereSampler stream over n_sims draws. 0 ok, -1 invalid_argument, -2 runtime_error.
int cer_sample(uint64_t seed, int n, double radius, double min_sep, uint64_t max_attempts,
               size_t n_sims, double* pos_out) {
  try {
    UniformSphereSampler::Options o;
    o.radius = static_cast<Real>(radius);
    o.min_separation = static_cast<Real>(min_sep);
    o.max_attempts_per_atom = max_attempts;
    UniformSphereSampler smp(o, seed);
    Molecule mol;
    for (int i = 0; i < n; ++i) mol.atoms.push_back(Atom{"X", 1, 1});
    State s;
    for (size_t k = 0; k < n_sims; ++k) {
      smp.sample(mol, s);
      for (int i = 0; i < n; ++i) {
        pos_out[(k * n + i) * 3] = s.positions[i].x;
        pos_out[(k * n + i) * 3 + 1] = s.positions[i].y;
        pos_out[(k * n + i) * 3 + 2] = s.positions[i].z;
      }
    }
  } catch (const std::invalid_argument&) {
    return -1;
  } catch (const std::runtime_error&) {
    return -2;
  }
  return 0;
}

// One run per sim with make_integrator(RK45) (Hairer start).
// stop_rule 0: run_to_convergence (driver.cpp).  stop_rule 1: the engine's
// momentum-change / t_max rule, driven through Integrator::step with the dt cap
// min(max_dt, t_max - t) and FP64 t, then the same redistribution as driver.cpp.
// status: 0 ok-pe, 1 ok-momentum, 2 ok-tmax, 3 not converged, 4 runtime_error.
int cer_run(int n, const double* m, const double* q, double k, double rtol, double atol,
            int stop_rule, double pe_stop, int redistribute, double max_dt, uint64_t max_steps,
            double t_max, double mom_tol, const double* pos, size_t n_sims, double* momenta,
            uint32_t* steps, uint8_t* status, double* t_final) {
  const Molecule mol = mk(n, m, q);
  const CoulombForce force(static_cast<Real>(k));
  for (size_t sim = 0; sim < n_sims; ++sim) {
    State s = st(n, pos + sim * n * 3);
    auto integ = make_integrator(IntegratorKind::RK45,
                                 IntegratorOptions{static_cast<Real>(rtol), static_cast<Real>(atol)});
    uint8_t code = 3;
    uint32_t nsteps = 0;
    double t = 0.0;
    try {
      if (stop_rule == 0) {
        RunConfig cfg;
        cfg.pe_stop_fraction = static_cast<Real>(pe_stop);
        cfg.redistribute_energy = redistribute != 0;
        cfg.max_dt = static_cast<Real>(max_dt);
        cfg.max_steps = max_steps;
        const RunResult r = run_to_convergence(mol, force, *integ, s, cfg);
        code = (r.converged || r.steps == 0) ? 0 : 3;
        nsteps = static_cast<uint32_t>(r.steps);
        t = r.t_final;
      } else {
        const Real e0 = force.potential_energy(mol, s) + CoulombForce::kinetic_energy(mol, s);
        const Real tol2 = static_cast<Real>(mom_tol * mom_tol);
        for (;;) {
          if (nsteps >= max_steps) break;
          const double rem = t_max - t;
          const bool by_time = rem <= max_dt;
          const Real cap = static_cast<Real>(by_time ? rem : max_dt);
          const std::vector<Vec3> vold = s.velocities;
          const Real dt = integ->step(mol, force, s, cap);
          ++nsteps;
          const bool final_step = by_time && (dt >= cap || t + static_cast<double>(dt) >= t_max);
          t = final_step ? t_max : t + static_cast<double>(dt);
          bool settled = !final_step;  // not tested on the step landing on t_max
          for (int i = 0; i < n && settled; ++i) {
            const Vec3 d = s.velocities[i] - vold[i];
            settled = dot(d, d) < tol2 * dot(s.velocities[i], s.velocities[i]);
          }
          if (settled) {
            code = 1;
            break;
          }
          if (final_step) {
            code = 2;
            break;
          }
        }
        const Real ke = CoulombForce::kinetic_energy(mol, s);
        if (redistribute && ke > Real{0}) {
          const Real sc = std::sqrt(e0 / ke);
          for (auto& v : s.velocities) v = v * sc;
        }
      }
    } catch (const std::runtime_error&) {
      code = 4;
    }
    for (int i = 0; i < n; ++i) {
      const Real mi = mol.atoms[i].mass;
      momenta[(sim * n + i) * 3] = s.velocities[i].x * mi;
      momenta[(sim * n + i) * 3 + 1] = s.velocities[i].y * mi;
      momenta[(sim * n + i) * 3 + 2] = s.velocities[i].z * mi;
    }
    steps[sim] = nsteps;
    status[sim] = code;
    if (t_final) t_final[sim] = t;
  }
  return 0;
}

}  // extern "C"
