// The reference generator's batch loop (examples/generate_dataset.cpp:168-227)
// written against the C++ mirror: host sampler with the reference's stream, then
// BatchedRK45::run.  Prints lane 0..k-1 momenta and steps as text for the test.
#include <cstdio>
#include <cstdlib>

#include "coulomb_gpu.hpp"

int main(int argc, char** argv) {
  using namespace coulomb::gpu;
  const std::size_t lanes = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 64;
  Molecule mol;
  mol.atoms.push_back(atom_from_amu("C", 12.011, 1.0));
  for (int i = 0; i < 4; ++i) mol.atoms.push_back(atom_from_amu("H", 1.008, 1.0));
  UniformSphereSampler sampler({4.0, 0.25, 10000}, 49188291);
  BatchedRK45 bi(mol, 1.0, lanes, 1e-4, 1e-7);
  for (std::size_t lane = 0; lane < lanes; ++lane) bi.set_lane_geometry(lane, sampler.sample(mol));
  RunConfig cfg;
  cfg.pe_stop_fraction = 1e-5;
  try {
    const auto res = bi.run(cfg);
    for (std::size_t s = 0; s < lanes; ++s) {
      std::printf("%zu %d", res.steps[s], static_cast<int>(res.converged[s]));
      for (std::size_t q = 0; q < mol.size() * 3; ++q)
        std::printf(" %.17g", res.momenta[s * mol.size() * 3 + q]);
      std::printf("\n");
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "run failed: %s\n", e.what());
    return 1;
  }
  try {
    BatchedRK45 bad(mol, 1.0, 0, 1e-4, 1e-7);  // zero lanes: invalid_argument
    return 2;
  } catch (const std::invalid_argument&) {
  }
  return 0;
}
