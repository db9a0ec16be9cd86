// The reference generator's batch loop (examples/generate_dataset.cpp:159-227)
// against include/coulomb_gpu.hpp: the same calls in the same order -- a
// UniformSphereSampler filling a State per lane through Sampler::sample(mol,
// State&), BatchedRK45<float>::set_lane_geometry(lane, state.positions),
// run(RunConfig) -- with the reference's record layout.  `lanes` comes from the
// command line instead of the host SIMD width (hn::Lanes), the only change.
//   generator_loop <lanes> <sims> <out.bin>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "coulomb_gpu.hpp"

using namespace coulomb::gpu;

namespace {

Molecule make_methane() {
  Molecule m;
  m.atoms.push_back(atom_from_amu("C", 12.011, 1.0));
  for (int i = 0; i < 4; ++i) m.atoms.push_back(atom_from_amu("H", 1.008, 1.0));
  return m;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: generator_loop <lanes> <sims> <out.bin>\n");
    return 2;
  }
  const std::size_t lanes = std::strtoull(argv[1], nullptr, 10);
  const std::size_t n_sims = std::strtoull(argv[2], nullptr, 10);
  const double rtol = 1e-4, atol = 1e-7, pe_stop = 1e-5;
  const Molecule mol = make_methane();
  const std::size_t n = mol.size();

  UniformSphereSampler sampler({.radius = static_cast<Real>(4.0),
                                .min_separation = static_cast<Real>(0.25)},
                               49188291);
  BatchedRK45<float> bi(mol, Real{1.0}, lanes, static_cast<float>(rtol),
                        static_cast<float>(atol));
  RunConfig cfg;
  cfg.pe_stop_fraction = static_cast<Real>(pe_stop);

  std::ofstream out(argv[3], std::ios::binary);
  if (!out) return 1;
  const std::size_t n_batches = (n_sims + lanes - 1) / lanes;
  std::vector<State> lane_geoms(lanes);
  std::vector<float> pos_buf(n * 3), mom_buf(n * 3);
  std::size_t written = 0;
  for (std::size_t b = 0; b < n_batches; ++b) {
    for (std::size_t lane = 0; lane < lanes; ++lane) {
      sampler.sample(mol, lane_geoms[lane]);
      bi.set_lane_geometry(lane, lane_geoms[lane].positions);
    }
    const auto res = bi.run(cfg);
    for (std::size_t lane = 0; lane < lanes && written < n_sims; ++lane, ++written) {
      for (std::size_t i = 0; i < n; ++i) {
        const Vec3& p = lane_geoms[lane].positions[i];
        pos_buf[i * 3 + 0] = static_cast<float>(p.x);
        pos_buf[i * 3 + 1] = static_cast<float>(p.y);
        pos_buf[i * 3 + 2] = static_cast<float>(p.z);
        for (int c = 0; c < 3; ++c)
          mom_buf[i * 3 + c] = static_cast<float>(res.momenta[(lane * n + i) * 3 + c]);
      }
      const std::uint32_t steps = static_cast<std::uint32_t>(res.steps[lane]);
      const std::uint8_t converged = static_cast<std::uint8_t>(res.converged[lane] ? 1 : 0);
      out.write(reinterpret_cast<const char*>(pos_buf.data()),
                static_cast<std::streamsize>(pos_buf.size() * sizeof(float)));
      out.write(reinterpret_cast<const char*>(mom_buf.data()),
                static_cast<std::streamsize>(mom_buf.size() * sizeof(float)));
      out.write(reinterpret_cast<const char*>(&steps), sizeof(steps));
      out.write(reinterpret_cast<const char*>(&converged), sizeof(converged));
    }
  }
  return out.good() ? 0 : 1;
}
