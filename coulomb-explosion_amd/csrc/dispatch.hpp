// Internal interface between the engine (engine.hip) and the per-precision
// translation units that instantiate the integrator kernels
// (dispatch_f32.hip, dispatch_f64.hip), so the two compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace ce {

constexpr int kMaxSpecies = kMaxSegments;

struct SpeciesConst {
  int n;
  double masses[10];
  double charges[10];  // fixed charges (per-sim charges come from RunArgs.qsim)
};

struct LaunchReq {
  RunArgs args;                    // segments filled in by the engine
  SpeciesConst species[kMaxSpecies];
  double coulomb_k;
  bool per_sim_charges;
  int block_threads;               // 0 = 256
  int blocks_per_cu;               // 0 = occupancy of the kernel
  int n_cus;
  hipStream_t stream;
};

struct LaunchInfo {
  int grid = 0;
  int block = 0;
  int launches = 0;
};

// Launches the integration of every segment in req.args.  Returns hipSuccess or
// the launch error; `bad_n` is set when a molecule size has no kernel.
hipError_t launch_dp5_f32(const LaunchReq& req, LaunchInfo* info, bool* bad_n);
hipError_t launch_dp5_f64(const LaunchReq& req, LaunchInfo* info, bool* bad_n);

template <class T, int N>
MolConst<T, N> make_mol_const(const SpeciesConst& s, double coulomb_k) {
  MolConst<T, N> mc{};
  for (int i = 0; i < N; ++i) {
    mc.im[i] = static_cast<T>(1.0 / s.masses[i]);
    mc.md[i] = s.masses[i];
    for (int j = i + 1; j < N; ++j) {
      const int p = pair_index<N>(i, j);
      mc.pcd[p] = coulomb_k * s.charges[i] * s.charges[j];
      mc.pc[p] = static_cast<T>(mc.pcd[p]);
    }
  }
  return mc;
}

// Single-size launch of segment `s` (its own queue slot).
template <class T, int N, bool PS>
hipError_t launch_one(const LaunchReq& req, int s, LaunchInfo* info) {
  const auto kern = dp5_persistent<T, N, PS>;
  const int block = req.block_threads > 0 ? req.block_threads : 256;
  int per_cu = req.blocks_per_cu;
  if (per_cu <= 0) {
    int occ = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, 0);
    if (e != hipSuccess) return e;
    per_cu = occ > 0 ? occ : 1;
  }
  RunArgs a = req.args;
  a.seg[0] = req.args.seg[s];
  a.n_segments = 1;
  a.queue_slot = s;
  const int grid = per_cu * req.n_cus;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, req.stream, a,
                     make_mol_const<T, N>(req.species[s], req.coulomb_k));
  info->grid = grid;
  info->block = block;
  info->launches += 1;
  return hipGetLastError();
}

template <class T, bool PS>
hipError_t launch_by_n(const LaunchReq& req, int s, LaunchInfo* info, bool* bad_n) {
  switch (req.species[s].n) {
    case 1: return launch_one<T, 1, PS>(req, s, info);
    case 2: return launch_one<T, 2, PS>(req, s, info);
    case 3: return launch_one<T, 3, PS>(req, s, info);
    case 4: return launch_one<T, 4, PS>(req, s, info);
    case 5: return launch_one<T, 5, PS>(req, s, info);
    case 6: return launch_one<T, 6, PS>(req, s, info);
    case 7: return launch_one<T, 7, PS>(req, s, info);
    case 8: return launch_one<T, 8, PS>(req, s, info);
    case 9: return launch_one<T, 9, PS>(req, s, info);
    case 10: return launch_one<T, 10, PS>(req, s, info);
    default: *bad_n = true; return hipSuccess;
  }
}

// The mixed launch compiled for sizes (2, 3, 5, 8) in that order; any other
// combination runs one launch per segment on the same stream.
template <class T, bool PS>
hipError_t launch_all(const LaunchReq& req, LaunchInfo* info, bool* bad_n) {
  const RunArgs& a = req.args;
  const bool mixed_2358 = a.n_segments == 4 && req.species[0].n == 2 && req.species[1].n == 3 &&
                          req.species[2].n == 5 && req.species[3].n == 8;
  if (mixed_2358) {
    using MC = MixedConst<T, 2, 3, 5, 8>;
    const auto kern = dp5_mixed<T, 2, 3, 5, 8, PS>;
    const int block = req.block_threads > 0 ? req.block_threads : 256;
    int per_cu = req.blocks_per_cu;
    if (per_cu <= 0) {
      int occ = 0;
      const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, 0);
      if (e != hipSuccess) return e;
      per_cu = occ > 0 ? occ : 1;
    }
    MC mc{make_mol_const<T, 2>(req.species[0], req.coulomb_k),
          make_mol_const<T, 3>(req.species[1], req.coulomb_k),
          make_mol_const<T, 5>(req.species[2], req.coulomb_k),
          make_mol_const<T, 8>(req.species[3], req.coulomb_k)};
    const int grid = per_cu * req.n_cus;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, req.stream, a, mc);
    info->grid = grid;
    info->block = block;
    info->launches = 1;
    return hipGetLastError();
  }
  for (int s = 0; s < a.n_segments; ++s) {
    if (a.seg[s].count == 0) continue;
    const hipError_t e = launch_by_n<T, PS>(req, s, info, bad_n);
    if (e != hipSuccess || *bad_n) return e;
  }
  return hipSuccess;
}

}  // namespace ce
