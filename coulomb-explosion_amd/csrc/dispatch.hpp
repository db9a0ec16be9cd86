// Internal interface between the engine (engine.hip) and the per-precision
// translation units that instantiate the integrator kernels
// (dispatch_f32.hip, dispatch_f64.hip), so the two compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "coulomb_explosion.h"
#include "hip_check.hpp"
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
  int mixed_launch;                // CE_MIXED_* for sizes (2, 3, 5, 8)
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

// True when every pair of the species has the same constant k q_i q_j, finite
// and non-zero (kUniform launches).
inline bool uniform_pairs(const SpeciesConst& s, double coulomb_k) {
  if (s.n < 2) return false;
  const double c = coulomb_k * s.charges[0] * s.charges[1];
  if (!(c != 0.0) || !__builtin_isfinite(c)) return false;
  for (int i = 0; i < s.n; ++i)
    for (int j = i + 1; j < s.n; ++j)
      if (coulomb_k * s.charges[i] * s.charges[j] != c) return false;
  return true;
}

// Kernel constants; a kUniform launch stores c / m_i in im (forces are summed
// without the common pair constant c = pc[0]).
template <class T, int N>
MolConst<T, N> make_mol_const(const SpeciesConst& s, double coulomb_k, int cm) {
  MolConst<T, N> mc{};
  for (int i = 0; i < N; ++i) {
    mc.md[i] = s.masses[i];
    for (int j = i + 1; j < N; ++j) {
      const int p = pair_index<N>(i, j);
      mc.pcd[p] = coulomb_k * s.charges[i] * s.charges[j];
      mc.pc[p] = static_cast<T>(mc.pcd[p]);
    }
  }
  const double c = cm == kUniform ? mc.pcd[0] : 1.0;
  for (int i = 0; i < N; ++i) mc.im[i] = static_cast<T>(c / s.masses[i]);
  return mc;
}

// Single-size launch of segment `s` (its own queue slot).
template <class T, int N, int CM>
hipError_t launch_one(const LaunchReq& req, int s, LaunchInfo* info) {
  const auto kern = dp5_persistent<T, N, CM>;
  const int block = req.block_threads > 0 ? req.block_threads : 256;
  int per_cu = req.blocks_per_cu;
  if (per_cu <= 0) {
    int occ = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, 0);
    if (e != hipSuccess) {
      ce::clear_stale_error();
      return e;
    }
    per_cu = occ > 0 ? occ : 1;
  }
  RunArgs a = req.args;
  a.seg[0] = req.args.seg[s];
  a.n_segments = 1;
  a.queue_slot = s;
  const int grid = per_cu * req.n_cus;
  ce::clear_stale_error();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, req.stream, a,
                     make_mol_const<T, N>(req.species[s], req.coulomb_k, CM));
  info->grid = grid;
  info->block = block;
  info->launches += 1;
  return hipGetLastError();
}

template <class T, int CM>
hipError_t launch_by_n(const LaunchReq& req, int s, LaunchInfo* info, bool* bad_n) {
  switch (req.species[s].n) {
    case 1: return launch_one<T, 1, CM>(req, s, info);
    case 2: return launch_one<T, 2, CM>(req, s, info);
    case 3: return launch_one<T, 3, CM>(req, s, info);
    case 4: return launch_one<T, 4, CM>(req, s, info);
    case 5: return launch_one<T, 5, CM>(req, s, info);
    case 6: return launch_one<T, 6, CM>(req, s, info);
    case 7: return launch_one<T, 7, CM>(req, s, info);
    case 8: return launch_one<T, 8, CM>(req, s, info);
    case 9: return launch_one<T, 9, CM>(req, s, info);
    case 10: return launch_one<T, 10, CM>(req, s, info);
    default: *bad_n = true; return hipSuccess;
  }
}

// One launch of the kernel compiled for sizes (2, 3, 5, 8), segments in that order.
template <class T, int CM>
hipError_t launch_mixed(const LaunchReq& req, LaunchInfo* info) {
  using MC = MixedConst<T, 2, 3, 5, 8>;
  const auto kern = dp5_mixed<T, 2, 3, 5, 8, CM>;
  const int block = req.block_threads > 0 ? req.block_threads : 256;
  int per_cu = req.blocks_per_cu;
  if (per_cu <= 0) {
    int occ = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, 0);
    if (e != hipSuccess) {
      ce::clear_stale_error();
      return e;
    }
    per_cu = occ > 0 ? occ : 1;
  }
  MC mc{make_mol_const<T, 2>(req.species[0], req.coulomb_k, CM),
        make_mol_const<T, 3>(req.species[1], req.coulomb_k, CM),
        make_mol_const<T, 5>(req.species[2], req.coulomb_k, CM),
        make_mol_const<T, 8>(req.species[3], req.coulomb_k, CM)};
  const int grid = per_cu * req.n_cus;
  ce::clear_stale_error();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, req.stream, req.args, mc);
  info->grid = grid;
  info->block = block;
  info->launches = 1;
  return hipGetLastError();
}

// The mixed launch compiled for sizes (2, 3, 5, 8) in that order (as
// req.mixed_launch selects); any other combination runs one launch per segment on
// the same stream.  Without per-sim
// charges and with UNIFORM_OK (the f32 path), a launch whose species all have a
// single pair constant runs the kUniform kernel.
template <class T, bool PS, bool UNIFORM_OK>
hipError_t launch_all(const LaunchReq& req, LaunchInfo* info, bool* bad_n) {
  const RunArgs& a = req.args;
  bool mixed_2358 = a.n_segments == 4 && req.species[0].n == 2 && req.species[1].n == 3 &&
                    req.species[2].n == 5 && req.species[3].n == 8;
  if (mixed_2358 && req.mixed_launch != CE_MIXED_ONE_LAUNCH) {
    // Per-species launches give the small molecules their own occupancy (the one
    // launch carries the N = 8 segment's registers) at the cost of a tail per launch.
    bool large = true;
    for (int s = 0; s < 4; ++s) large = large && a.seg[s].count >= (uint64_t(1) << 19);
    mixed_2358 = req.mixed_launch == CE_MIXED_AUTO && !large;
  }
  if (mixed_2358) {
    if constexpr (!PS && UNIFORM_OK) {
      bool all_uniform = true;
      for (int s = 0; s < 4; ++s) all_uniform = all_uniform && uniform_pairs(req.species[s], req.coulomb_k);
      if (all_uniform) return launch_mixed<T, kUniform>(req, info);
    }
    return launch_mixed<T, PS ? kPerSim : kPairTable>(req, info);
  }
  for (int s = 0; s < a.n_segments; ++s) {
    if (a.seg[s].count == 0) continue;
    hipError_t e;
    if constexpr (PS) {
      e = launch_by_n<T, kPerSim>(req, s, info, bad_n);
    } else if constexpr (UNIFORM_OK) {
      e = uniform_pairs(req.species[s], req.coulomb_k)
              ? launch_by_n<T, kUniform>(req, s, info, bad_n)
              : launch_by_n<T, kPairTable>(req, s, info, bad_n);
    } else {
      e = launch_by_n<T, kPairTable>(req, s, info, bad_n);
    }
    if (e != hipSuccess || *bad_n) return e;
  }
  return hipSuccess;
}

}  // namespace ce
