// Engine state shared by the C-ABI translation units (engine.hip, closure.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "coulomb_explosion.h"
#include "dispatch.hpp"
#include "kernels.hpp"

namespace ce {

struct Species {
  int n = 0;
  double masses[CE_MAX_ATOMS] = {};
  double charges[CE_MAX_ATOMS] = {};
  double radius = 0, min_sep = 0;
  int levels = 0;      // per-sim charge levels (0 = fixed charges)
  int max_total = 0;
  uint64_t capacity = 0;  // sims this species can hold
  uint64_t count = 0;     // sims loaded for the next run
};


// Sets the thread's last-error message and returns `code`.
int set_error(int code, const std::string& msg);

}  // namespace ce

struct ce_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t k0 = nullptr, k1 = nullptr, s0 = nullptr, s1 = nullptr, p0 = nullptr, p1 = nullptr;
  int n_species = 0;
  ce::Species sp[CE_MAX_SPECIES];
  bool mixed = false;        // created by ce_engine_create_mixed (fixed species counts)
  double coulomb_k = 1.0;
  double rtol = 0, atol = 0;
  size_t total_pos = 0;      // sum capacity * 3n (doubles)
  size_t total_q = 0;        // sum capacity * n
  size_t capacity = 0;       // sum of species capacities
  bool sampled = false;
  bool have_charges = false; // qsim holds per-sim charges for the loaded sims
  // device buffers
  double* geom = nullptr;
  uint8_t* gstat = nullptr;
  float* qsim = nullptr;
  unsigned long long* ctr = nullptr;
  double* mom = nullptr;
  float* mom32 = nullptr;
  uint32_t* steps = nullptr;
  uint32_t* attempts = nullptr;
  uint8_t* status = nullptr;
  double* tfin = nullptr;
  double* eerr = nullptr;
  double* trace = nullptr;
  uint32_t trace_cap = 0;
  bool energy_diag = false;  // compute eerr on the next runs
  bool eerr_valid = false;   // the last run computed eerr
  uint8_t* records = nullptr;
  size_t records_cap = 0;
  // queued record copies (ce_engine_pack_records_async): two device buffers
  // used in turn, packed on `stream`, copied to the host on `copy_stream`
  hipStream_t copy_stream = nullptr;
  uint8_t* rec_async[2] = {nullptr, nullptr};
  size_t rec_async_cap[2] = {0, 0};
  hipEvent_t rec_packed[2] = {nullptr, nullptr}, rec_copied[2] = {nullptr, nullptr};
  int rec_slot = 0;
  // integration state at the stop (ce_engine_enable_state): positions and
  // velocities in the run's precision, next step size, E0
  void* state_x = nullptr;
  void* state_v = nullptr;
  double* state_h = nullptr;
  double* state_e0 = nullptr;
  int state_prec = -1;        // precision of the run that wrote the state, -1 none
  uint32_t* list = nullptr;   // sim list of ce_engine_run_subset [capacity]
  // captured batch (ce_engine_capture): sampler + counter reset + integrator
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint64_t* first_dev = nullptr;  // first sim index of the replayed batch
  ce_run_config graph_cfg{};
  size_t graph_sims = 0;
  bool graph_eerr = false;  // the captured integrator computes eerr
  ce::LaunchInfo graph_info{};
  // launch
  ce_launch_config launch = {0, 0, 4, 64};
  int n_cus = 0;
  ce_run_stats stats = {};
  bool have_results = false;

  size_t loaded() const {
    size_t t = 0;
    for (int s = 0; s < n_species; ++s) t += sp[s].count;
    return t;
  }
  // start / offsets of species s for the loaded counts (single species: 0)
  void offsets(int s, uint64_t* start, uint64_t* pos_off, uint64_t* q_off) const {
    uint64_t a = 0, b = 0, c = 0;
    for (int r = 0; r < s; ++r) {
      a += sp[r].capacity;
      b += sp[r].capacity * 3 * sp[r].n;
      c += sp[r].capacity * sp[r].n;
    }
    *start = a;
    *pos_off = b;
    *q_off = c;
  }
};

namespace ce {

// Runs the loaded sims (or, single-species engines, the segment-local sims of
// the device list `list` of n_list entries) with `cfg`; results stay on the
// device.
int engine_run(ce_engine* e, const ce_run_config* cfg, const uint32_t* list, size_t n_list);

}  // namespace ce
