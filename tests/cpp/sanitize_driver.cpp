// Exercises the library's host code under AddressSanitizer and
// UndefinedBehaviorSanitizer (host code only: the device code is not
// instrumented).  Part 1 runs without a device: argument checks, error paths,
// the host sampler.  Part 2, when a device is visible and "device" is passed:
// engine lifetimes, every copy-in / copy-out path, saved state, subsets,
// captured graphs, queued records into pinned memory, mixed engines, the MLP.
// Prints "sanitize driver ok" and exits 0; a sanitizer report exits non-zero.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "coulomb_explosion.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s (%s)\n", what, ce_last_error());
    ++failures;
  }
}

void host_part() {
  ce_run_config cfg;
  ce_run_config_default(&cfg);
  expect(cfg.max_steps == 1000000 && cfg.pe_stop_fraction == 1e-9, "config defaults");
  expect(ce_record_bytes(5) == 125 && ce_record_bytes(0) == 0, "record bytes");
  // argument checks before any device access
  const double m[5] = {21874.66, 1837.15, 1837.15, 1837.15, 1837.15}, q[5] = {1, 1, 1, 1, 1};
  ce_molecule_desc mol{5, 0, m, q, 1.0};
  ce_tolerances tol{1e-6, 1e-9};
  ce_engine* e = nullptr;
  expect(ce_engine_create(nullptr, &tol, 10, &e) == CE_ERR_INVALID_ARGUMENT, "null molecule");
  ce_molecule_desc bad = mol;
  bad.n_atoms = 11;
  expect(ce_engine_create(&bad, &tol, 10, &e) == CE_ERR_INVALID_ARGUMENT, "too many atoms");
  const double neg[5] = {1, -1, 1, 1, 1};
  bad = mol;
  bad.masses = neg;
  expect(ce_engine_create(&bad, &tol, 10, &e) == CE_ERR_INVALID_ARGUMENT, "negative mass");
  ce_tolerances badtol{0.0, 1e-9};
  expect(ce_engine_create(&mol, &badtol, 10, &e) == CE_ERR_INVALID_ARGUMENT, "rtol 0");
  expect(ce_engine_create(&mol, &tol, 0, &e) == CE_ERR_INVALID_ARGUMENT, "capacity 0");
  expect(ce_engine_run(nullptr, &cfg, nullptr) == CE_ERR_INVALID_ARGUMENT, "run null");
  expect(ce_run_device_noalloc(nullptr, nullptr, nullptr, nullptr, 1, nullptr, nullptr) ==
             CE_ERR_INVALID_ARGUMENT, "noalloc null");
  expect(std::strlen(ce_last_error()) > 0, "error message");
  // the host sampler: the same stream after a reseed, error paths
  ce_host_sampler* hs = nullptr;
  expect(ce_host_sampler_create(7, -1.0, 0.25, 100, &hs) == CE_ERR_INVALID_ARGUMENT, "radius");
  expect(ce_host_sampler_create(7, 2.5, 1.0, 10000, &hs) == CE_OK, "sampler create");
  std::vector<double> a(1000 * 15), b(1000 * 15);
  expect(ce_host_sampler_sample(hs, 5, 1000, a.data()) == CE_OK, "sample");
  expect(ce_host_sampler_reseed(hs, 7) == CE_OK, "reseed");
  expect(ce_host_sampler_sample(hs, 5, 1000, b.data()) == CE_OK, "sample again");
  expect(a == b, "reseeded stream");
  std::vector<double> c(30);
  ce_host_sampler* tight = nullptr;
  expect(ce_host_sampler_create(1, 1.0, 1.5, 100, &tight) == CE_OK, "tight create");
  expect(ce_host_sampler_sample(tight, 10, 1, c.data()) == CE_ERR_RUNTIME, "infeasible");
  ce_host_sampler_destroy(tight);
  ce_host_sampler_destroy(hs);
}

void device_part() {
  const double m[5] = {21874.66, 1837.15, 1837.15, 1837.15, 1837.15}, q[5] = {1, 1, 1, 1, 1};
  ce_molecule_desc mol{5, 0, m, q, 1.0};
  ce_tolerances tol{1e-6, 1e-9};
  ce_run_config cfg;
  ce_run_config_default(&cfg);
  cfg.stop_rule = CE_STOP_MOMENTUM_OR_TMAX;
  cfg.h0_mode = CE_H0_FIXED;
  const size_t ns = 5000;
  ce_engine* e = nullptr;
  expect(ce_engine_create(&mol, &tol, ns, &e) == CE_OK, "create");
  expect(ce_engine_sample(e, 11, 0, ns, 2.5, 1.0, 10000) == CE_OK, "sample");
  expect(ce_engine_enable_trace(e, 64) == CE_OK, "trace");
  expect(ce_engine_enable_energy_diagnostic(e, 1) == CE_OK, "energy");
  expect(ce_engine_enable_state(e, 1) == CE_OK, "state");
  std::vector<double> mom(ns * 15), tf(ns), ee(ns), tr(ns * 64), geom(ns * 15);
  std::vector<float> mom32(ns * 15);
  std::vector<uint32_t> steps(ns), att(ns);
  std::vector<uint8_t> st(ns);
  ce_results out{mom.data(), mom32.data(), steps.data(), att.data(), st.data(), tf.data(),
                 ee.data(), tr.data(), 64, 0};
  expect(ce_engine_run(e, &cfg, &out) == CE_OK, "run");
  // every sim integrated to its stop (a run that did nothing leaves zeros)
  const auto all_ran = [&](const char* what) {
    bool ok = true;
    for (size_t s = 0; s < ns; ++s)
      ok = ok && steps[s] > 0 && att[s] >= steps[s] && st[s] <= CE_SIM_OK_TMAX && mom[s * 15] != 0.0;
    expect(ok, what);
  };
  all_ran("run results");
  expect(ce_engine_get_geometries(e, geom.data(), ns) == CE_OK, "geometries");
  std::vector<double> x(ns * 15), v(ns * 15), h(ns), e0(ns);
  expect(ce_engine_get_state(e, x.data(), v.data(), h.data(), e0.data(), ns) == CE_OK, "get state");
  expect(ce_engine_set_state(e, x.data(), v.data(), h.data(), e0.data(), ns,
                             CE_PRECISION_F32) == CE_OK, "set state");
  ce_run_config res = cfg;
  res.resume = 1;
  res.t_max = 2e4;
  expect(ce_engine_run(e, &res, &out) == CE_OK, "resume");
  const uint64_t sub[4] = {0, 7, 4999, 123};
  expect(ce_engine_run_subset(e, &res, sub, 4, nullptr) == CE_OK, "subset");
  uint64_t bad_sub[1] = {ns};
  expect(ce_engine_run_subset(e, &res, bad_sub, 1, nullptr) == CE_ERR_INVALID_ARGUMENT, "subset range");
  expect(ce_engine_enable_trace(e, 0) == CE_OK && ce_engine_enable_energy_diagnostic(e, 0) == CE_OK,
         "disable");
  out.trace_h = nullptr;
  out.trace_cap = 0;
  out.energy_err = nullptr;
  expect(ce_engine_run(e, &cfg, &out) == CE_OK, "run again");
  all_ran("second run results");
  ce_run_stats st0{};
  expect(ce_engine_stats(e, &st0) == CE_OK && st0.attempted_steps >= ns, "attempted steps");
  std::vector<uint8_t> rec(ns * 125);
  expect(ce_engine_pack_records(e, 0, ns, rec.data()) == CE_OK, "records");
  void* pin[2] = {nullptr, nullptr};
  for (auto& p : pin) expect(ce_host_alloc(ns * 125, &p) == CE_OK, "pinned alloc");
  expect(ce_engine_pack_records_async(e, 0, ns, static_cast<uint8_t*>(pin[0])) == CE_OK, "async 1");
  expect(ce_engine_pack_records_async(e, 100, 200, static_cast<uint8_t*>(pin[1])) == CE_OK, "async 2");
  expect(ce_engine_records_wait(e) == CE_OK, "records wait");
  expect(std::memcmp(pin[0], rec.data(), ns * 125) == 0, "async records");
  for (void* p : pin) ce_host_free(p);
  expect(ce_engine_capture(e, &cfg, 11, ns, 2.5, 1.0, 10000) == CE_OK, "capture");
  expect(ce_engine_replay(e, 0, &out) == CE_OK, "replay");
  all_ran("replay results");
  ce_launch_config lc{};
  expect(ce_engine_get_launch(e, &lc) == CE_OK && ce_engine_set_launch(e, &lc) == CE_OK, "launch");
  ce_run_stats stats{};
  expect(ce_engine_stats(e, &stats) == CE_OK && stats.sims == ns, "stats");
  expect(ce_engine_enable_state(e, 0) == CE_OK, "state off");
  ce_engine_destroy(e);
  // f64 path and a capacity the device cannot hold (allocation failure path)
  expect(ce_engine_create(&mol, &tol, 1000, &e) == CE_OK, "create 2");
  expect(ce_engine_sample(e, 3, 0, 1000, 2.5, 1.0, 10000) == CE_OK, "sample 2");
  cfg.precision = CE_PRECISION_F64;
  expect(ce_engine_run(e, &cfg, nullptr) == CE_OK, "f64 run");
  ce_engine_destroy(e);
  ce_engine* huge = nullptr;
  expect(ce_engine_create(&mol, &tol, size_t(1) << 40, &huge) != CE_OK && huge == nullptr, "huge");
  // mixed engine with per-sim charges
  const double mn2[2] = {25526.04, 25526.04}, mco2[3] = {21874.66, 29156.95, 29156.95};
  const double mc2h6[8] = {21874.66, 21874.66, 1837.15, 1837.15, 1837.15, 1837.15, 1837.15, 1837.15};
  ce_species_desc sp[4] = {{2, 2, mn2, nullptr, 2.76, 0.9, 500, 4, 0},
                           {3, 2, mco2, nullptr, 2.94, 0.9, 500, 6, 0},
                           {5, 2, m, nullptr, 3.21, 0.9, 500, 10, 0},
                           {8, 2, mc2h6, nullptr, 3.5, 0.9, 500, 16, 0}};
  expect(ce_engine_create_mixed(sp, 4, 1.0, &tol, &e) == CE_OK, "mixed create");
  expect(ce_engine_sample_species(e, 5, 0, 10000) == CE_OK, "mixed sample");
  cfg.precision = CE_PRECISION_F32;
  std::vector<double> mm(500 * 3 * (2 + 3 + 5 + 8)), qq(500 * (2 + 3 + 5 + 8));
  ce_results mo{};
  mo.momenta = mm.data();
  expect(ce_engine_run(e, &cfg, &mo) == CE_OK, "mixed run");
  expect(ce_engine_get_charges(e, qq.data(), 2000) == CE_OK, "mixed charges");
  ce_engine_destroy(e);
  // the closure network
  const int32_t widths[4] = {40, 64, 64, 15};
  ce_mlp* net = nullptr;
  expect(ce_mlp_create(widths, 4, 1, &net) == CE_OK, "mlp create");
  std::vector<float> xin(256 * 40, 0.5f), tgt(256 * 15, 0.1f), yout(256 * 15),
      grad(static_cast<size_t>(ce_mlp_param_count(net)));
  float loss = 0.0f;
  expect(ce_mlp_forward(net, xin.data(), 256, yout.data()) == CE_OK, "mlp forward");
  expect(ce_mlp_gradients(net, xin.data(), tgt.data(), 256, grad.data(), &loss) == CE_OK, "grad");
  expect(ce_mlp_adam_step(net, 1e-3f) == CE_OK, "adam");
  ce_mlp_destroy(net);
}

}  // namespace

int main(int argc, char** argv) {
  host_part();
  int ndev = 0;
  const bool want_device = argc > 1 && std::string(argv[1]) == "device";
  if (want_device) {
    expect(ce_device_count(&ndev) == CE_OK && ndev > 0, "a device is visible");
    if (ndev > 0) device_part();
  }
  if (failures) return 1;
  std::printf("sanitize driver ok (%s)\n", want_device ? "host and device paths" : "host paths");
  return 0;
}
