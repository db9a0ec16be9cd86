// The reference's strategy API (sampler, force, integrator, driver) through
// include/coulomb_gpu.hpp, written as a reference caller would write it.  Prints
// one "key values..." line per quantity for tests/test_gpu_cpp_mirror.py, which
// checks them against the oracle.
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <vector>

#include "coulomb_gpu.hpp"

using namespace coulomb::gpu;

namespace {

void print_vecs(const char* key, const std::vector<Vec3>& v) {
  std::printf("%s", key);
  for (const Vec3& a : v) std::printf(" %.17g %.17g %.17g", a.x, a.y, a.z);
  std::printf("\n");
}

// An integrator the driver knows nothing about (it only forwards), so
// run_to_convergence takes its generic loop.
class Forwarding final : public Integrator {
 public:
  explicit Forwarding(const IntegratorOptions& o) : inner_(o) {}
  Real step(const Molecule& m, const CoulombForce& f, State& s, Real dt) override {
    return inner_.step(m, f, s, dt);
  }
  std::string_view name() const override { return "forwarding"; }

 private:
  DormandPrince45 inner_;
};

void print_run(const char* key, const RunResult& r, const State& s, std::size_t observed,
               Real observed_t) {
  std::printf("%s %zu %.17g %.17g %.17g %.17g %.17g %.17g %d %zu %.17g\n", key, r.steps,
              r.t_final, r.energy_initial, r.pe_final, r.ke_before_redist, r.ke_after_redist,
              r.redistribution_scale, r.converged ? 1 : 0, observed, observed_t);
  const std::string vk = std::string(key) + "_vel";
  print_vecs(vk.c_str(), s.velocities);
}

}  // namespace

int main() {
  Molecule mol;
  const double masses[5] = {21874.66, 1837.15, 1837.15, 1837.15, 1837.15};
  const double charges[5] = {1, 2, 1, 1, 2};
  for (int i = 0; i < 5; ++i)
    mol.atoms.push_back(Atom{"X", static_cast<Real>(masses[i]), static_cast<Real>(charges[i])});

  std::unique_ptr<Sampler> sampler =
      std::make_unique<UniformSphereSampler>(UniformSphereSampler::Options{2.5, 1.0, 10000}, 12345);
  State start;
  sampler->sample(mol, start);
  std::printf("sampler_name %s\n", std::string(sampler->name()).c_str());
  print_vecs("geom", start.positions);
  print_vecs("vel0", start.velocities);

  const CoulombForce force(1.0);
  std::vector<Vec3> acc;
  force.accelerations(mol, start, acc);
  print_vecs("acc", acc);
  std::printf("pe %.17g\n", force.potential_energy(mol, start));
  std::printf("ke %.17g\n", CoulombForce::kinetic_energy(mol, start));

  const IntegratorOptions opts{1e-8, 1e-12};
  auto integ = make_integrator(IntegratorKind::RK45, opts);
  std::printf("integrator_name %s\n", std::string(integ->name()).c_str());
  State s = start;
  std::printf("dt");
  for (int k = 0; k < 12; ++k) std::printf(" %.17g", integ->step(mol, force, s, 1e12));
  std::printf("\n");
  print_vecs("step_vel", s.velocities);
  print_vecs("step_pos", s.positions);

  RunConfig cfg;
  cfg.pe_stop_fraction = 1e-6;
  for (int pass = 0; pass < 2; ++pass) {
    std::unique_ptr<Integrator> it = pass == 0 ? make_integrator(IntegratorKind::RK45, opts)
                                               : std::make_unique<Forwarding>(opts);
    State st = start;
    std::size_t observed = 0;
    Real t_obs = 0;
    const RunResult r = run_to_convergence(mol, force, *it, st, cfg,
                                           [&](std::size_t i, Real t0, Real dt) {
                                             if (i != observed || t0 != t_obs)
                                               std::printf("observer_out_of_order %zu\n", i);
                                             ++observed;
                                             t_obs = t0 + dt;
                                           });
    print_run(pass == 0 ? "run_engine" : "run_loop", r, st, observed, t_obs);
  }

  try {
    make_integrator(IntegratorKind::VelocityVerlet);
    std::printf("vv_throws 0\n");
  } catch (const std::invalid_argument&) {
    std::printf("vv_throws 1\n");
  }
  return 0;
}
