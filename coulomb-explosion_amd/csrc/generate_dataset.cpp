// Dataset generator on the MI355X engine: the counterpart of the reference's
// examples/generate_dataset.cpp (same flags and defaults, same binary records
// and .meta.json sidecar, readable by examples/to_parquet.py), driving the
// C-ABI batch by batch.  Sims are sampled on the device by default (counter-based,
// keyed by (seed, global sim index)); --sampler host reproduces the
// reference's std::mt19937_64 stream for a seed instead.
//
//   generate_dataset --out PATH [--sims N] [--seed S] [--radius R]
//       [--min-separation D] [--rtol T] [--atol A] [--pe-stop P]
//       [--stop-rule pe|tmax] [--t-max T] [--momentum-tol M] [--h0 H (0 = Hairer)]
//       [--batch B] [--sampler device|host] [--masses m1,m2,.. --charges q1,q2,..]
//       [--progress-every B] [--no-graph]
// Records go through two pinned host buffers: the device packs and copies batch
// b's records while it computes batch b + 1, and a writer thread writes them
// while it computes batch b + 2.  --out /dev/null measures the pipeline with the
// records in host memory (no sidecar is written then).
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "coulomb_explosion.h"
#include "record_writer.hpp"

namespace {

struct Args {
  std::size_t sims = 10'000'000;
  std::uint64_t seed = 49188291;
  double radius = 4.0;
  double min_separation = 0.25;
  double rtol = 1e-4;
  double atol = 1e-7;
  double pe_stop = 1e-5;
  std::string stop_rule = "pe";
  double t_max = 1e4;
  double momentum_tol = 1e-7;
  double h0 = 0.0;
  std::size_t batch = 10'000'000;
  std::string sampler = "device";
  std::vector<double> masses;
  std::vector<double> charges;
  std::vector<std::string> symbols;
  std::string out;
  std::size_t progress_every = 1;
  bool graph = true;  // replay full device-sampled batches from a captured HIP graph
};

std::vector<double> parse_list(const std::string& s) {
  std::vector<double> v;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, ',')) v.push_back(std::stod(item));
  return v;
}

[[noreturn]] void die(const std::string& msg) {
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::exit(EXIT_FAILURE);
}

void check(int rc, const char* what) {
  if (rc != CE_OK) die(std::string(what) + ": " + ce_last_error());
}

Args parse_args(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    const std::string f = argv[i];
    if (f == "--no-graph") {
      a.graph = false;
      continue;
    }
    if (i + 1 >= argc) die("unknown or incomplete argument: " + f);
    const std::string v = argv[++i];
    try {
      if (f == "--sims") a.sims = std::stoull(v);
      else if (f == "--seed") a.seed = std::stoull(v);
      else if (f == "--radius") a.radius = std::stod(v);
      else if (f == "--min-separation") a.min_separation = std::stod(v);
      else if (f == "--rtol") a.rtol = std::stod(v);
      else if (f == "--atol") a.atol = std::stod(v);
      else if (f == "--pe-stop") a.pe_stop = std::stod(v);
      else if (f == "--stop-rule") a.stop_rule = v;
      else if (f == "--t-max") a.t_max = std::stod(v);
      else if (f == "--momentum-tol") a.momentum_tol = std::stod(v);
      else if (f == "--h0") a.h0 = std::stod(v);
      else if (f == "--batch") a.batch = std::stoull(v);
      else if (f == "--sampler") a.sampler = v;
      else if (f == "--masses") a.masses = parse_list(v);
      else if (f == "--charges") a.charges = parse_list(v);
      else if (f == "--out") a.out = v;
      else if (f == "--progress-every") a.progress_every = std::stoull(v);
      else die("unknown or incomplete argument: " + f);
    } catch (const std::exception&) {  // std::stoull / std::stod on a malformed value
      die("bad value for " + f + ": " + v);
    }
  }
  if (a.out.empty()) die("--out is required (no default -- pick your own output location)");
  if (a.stop_rule != "pe" && a.stop_rule != "tmax") die("--stop-rule must be pe or tmax");
  if (a.sampler != "device" && a.sampler != "host") die("--sampler must be device or host");
  if (a.masses.empty()) {
    // CH4, singly ionised, the reference generator's molecule
    // (generate_dataset.cpp:56-66): C 12.011 amu, 4 x H 1.008 amu, all +1.
    const double amu = 1822.888486209;
    a.masses = {12.011 * amu, 1.008 * amu, 1.008 * amu, 1.008 * amu, 1.008 * amu};
    a.charges = {1, 1, 1, 1, 1};
    a.symbols = {"C", "H", "H", "H", "H"};
  } else {
    if (a.charges.size() != a.masses.size()) die("--charges must match --masses");
    for (std::size_t i = 0; i < a.masses.size(); ++i) a.symbols.push_back("X" + std::to_string(i));
  }
  if (a.batch == 0) die("--batch must be positive");
  return a;
}

void write_metadata(const Args& a, double wall, double rate, std::size_t n_failed) {
  std::ofstream out(a.out + ".meta.json");
  if (!out) die("cannot open metadata file: " + a.out + ".meta.json");
  out.precision(17);  // masses and tolerances read back exactly
  const double amu = 1822.888486209;
  const std::size_t n = a.masses.size();
  out << "{\n  \"symbols\": [";
  for (std::size_t i = 0; i < n; ++i) out << "\"" << a.symbols[i] << "\"" << (i + 1 < n ? ", " : "");
  out << "],\n  \"masses_amu\": [";
  for (std::size_t i = 0; i < n; ++i) out << a.masses[i] / amu << (i + 1 < n ? ", " : "");
  out << "],\n  \"charges\": [";
  for (std::size_t i = 0; i < n; ++i) out << a.charges[i] << (i + 1 < n ? ", " : "");
  out << "],\n"
      << "  \"n_atoms\": " << n << ",\n"
      << "  \"n_sims\": " << a.sims << ",\n"
      << "  \"seed\": " << a.seed << ",\n"
      << "  \"radius\": " << a.radius << ",\n"
      << "  \"min_separation\": " << a.min_separation << ",\n"
      << "  \"rtol\": " << a.rtol << ",\n"
      << "  \"atol\": " << a.atol << ",\n"
      << "  \"pe_stop\": " << a.pe_stop << ",\n"
      << "  \"lanes\": " << a.batch << ",\n"
      << "  \"n_failed_convergence\": " << n_failed << ",\n"
      << "  \"wall_seconds\": " << wall << ",\n"
      << "  \"sims_per_sec\": " << rate << ",\n"
      << "  \"engine\": \"mi355x-gfx950 per-lane DP5(4)\",\n"
      << "  \"sampler\": \"" << (a.sampler == "device" ? "philox4x32-10 uniform-sphere" : "mt19937_64 uniform-sphere") << "\",\n"
      << "  \"stop_rule\": \"" << a.stop_rule << "\",\n"
      << "  \"t_max\": " << a.t_max << ",\n"
      << "  \"momentum_tol\": " << a.momentum_tol << ",\n"
      << "  \"h0\": " << a.h0 << ",\n"
      << "  \"record_layout\": ["
      << "\"init_pos: n*3 float32 (x,y,z per atom, atom order as symbols)\", "
      << "\"final_momentum: n*3 float32 (m*v per atom, post energy-redistribution)\", "
      << "\"steps: uint32\", "
      << "\"converged: uint8\"]\n"
      << "}\n";
}

}  // namespace

int main(int argc, char** argv) {
  const Args a = parse_args(argc, argv);
  const int32_t n = static_cast<int32_t>(a.masses.size());
  const std::size_t batch = std::min(a.batch, a.sims);
  std::printf("dataset generator: n_sims=%zu n_atoms=%d batch=%zu rtol=%.0e atol=%.0e stop=%s\n",
              a.sims, n, batch, a.rtol, a.atol, a.stop_rule.c_str());

  ce_molecule_desc mol{n, 0, a.masses.data(), a.charges.data(), 1.0};
  ce_tolerances tol{a.rtol, a.atol};
  ce_engine* eng = nullptr;
  check(ce_engine_create(&mol, &tol, batch, &eng), "ce_engine_create");
  ce_host_sampler* hs = nullptr;
  if (a.sampler == "host")
    check(ce_host_sampler_create(a.seed, a.radius, a.min_separation, 10000, &hs),
          "ce_host_sampler_create");
  ce_run_config cfg;
  ce_run_config_default(&cfg);
  cfg.pe_stop_fraction = a.pe_stop;
  cfg.stop_rule = a.stop_rule == "pe" ? CE_STOP_PE_FRACTION : CE_STOP_MOMENTUM_OR_TMAX;
  cfg.t_max = a.t_max;
  cfg.momentum_tol = a.momentum_tol;
  if (a.h0 > 0.0) {
    cfg.h0_mode = CE_H0_FIXED;
    cfg.h0 = a.h0;
  }

  const std::size_t rb = ce_record_bytes(n);
  // Two pinned host buffers: batch b's records are copied into buffer b % 2 while
  // the device computes batch b + 1, then written while it computes batch b + 2.
  uint8_t* host[2] = {nullptr, nullptr};
  for (auto& h : host)
    check(ce_host_alloc(batch * rb, reinterpret_cast<void**>(&h)), "ce_host_alloc");
  std::vector<double> host_pos;
  if (hs) host_pos.resize(batch * 3 * n);
  std::unique_ptr<ce::RecordWriter> writer_owner;
  try {
    writer_owner = std::make_unique<ce::RecordWriter>(a.out);
  } catch (const std::exception& e) {
    die(e.what());
  }
  ce::RecordWriter& writer = *writer_owner;
  std::vector<uint8_t> status(batch);
  std::size_t done = 0, b = 0, prev_cnt = 0, n_failed = 0;
  const auto t0 = std::chrono::steady_clock::now();
  auto t_first = t0;
  // Full device-sampled batches replay one captured HIP graph (sampler +
  // integrator); a final partial batch and host-sampled batches run eagerly.
  const bool graph = !hs && a.graph && a.sims >= batch;
  if (graph)
    check(ce_engine_capture(eng, &cfg, a.seed, batch, a.radius, a.min_separation, 10000),
          "ce_engine_capture");
  while (done < a.sims) {
    const std::size_t cnt = std::min(batch, a.sims - done);
    ce_results res{};
    res.status = status.data();  // 1 byte per sim: the failures are counted from it
    if (graph && cnt == batch) {
      check(ce_engine_replay(eng, done, &res), "ce_engine_replay");
    } else {
      if (hs) {
        check(ce_host_sampler_sample(hs, n, cnt, host_pos.data()), "ce_host_sampler_sample");
        check(ce_engine_set_geometries(eng, host_pos.data(), cnt), "ce_engine_set_geometries");
      } else {
        check(ce_engine_sample(eng, a.seed, done, cnt, a.radius, a.min_separation, 10000),
              "ce_engine_sample");
      }
      check(ce_engine_run(eng, &cfg, &res), "ce_engine_run");
    }
    n_failed += static_cast<std::size_t>(
        std::count_if(status.begin(), status.begin() + static_cast<std::ptrdiff_t>(cnt),
                      [](uint8_t s) { return s > CE_SIM_OK_TMAX; }));
    if (b > 0) {  // batch b - 1's records are in host memory once its copy is done
      check(ce_engine_records_wait(eng), "ce_engine_records_wait");
      writer.submit(host[(b - 1) % 2], prev_cnt * rb);
    }
    if (b >= 2) writer.wait_written(b - 1);  // batch b - 2 has left buffer b % 2
    check(ce_engine_pack_records_async(eng, 0, cnt, host[b % 2]), "ce_engine_pack_records_async");
    done += cnt;
    prev_cnt = cnt;
    ++b;
    if (b == 1) t_first = std::chrono::steady_clock::now();
    if (b % a.progress_every == 0 || done == a.sims) {
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::fprintf(stderr, "\r%zu/%zu sims (%.0f sims/sec)   ", done, a.sims, done / el);
      std::fflush(stderr);
    }
  }
  check(ce_engine_records_wait(eng), "ce_engine_records_wait");
  writer.submit(host[(b - 1) % 2], prev_cnt * rb);
  try {
    writer.finish();
  } catch (const std::exception& e) {
    die(e.what());
  }
  std::fprintf(stderr, "\n");
  const auto t_end = std::chrono::steady_clock::now();
  const double wall = std::chrono::duration<double>(t_end - t0).count();
  const double rate = static_cast<double>(a.sims) / wall;
  std::printf("done: %zu sims in %.2fs (%.0f sims/sec), %zu failed to converge\n", a.sims, wall,
              rate, n_failed);
  if (b > 1) {
    // after the first batch (graph capture, first-touch of buffers): the pipeline's rate
    const double w = std::chrono::duration<double>(t_end - t_first).count();
    std::printf("steady: %zu sims in %.3fs after the first batch (%.0f sims/sec)\n",
                a.sims - std::min(batch, a.sims), w, (a.sims - std::min(batch, a.sims)) / w);
  }
  for (uint8_t* h : host) ce_host_free(h);
  if (a.out == "/dev/null") {
    std::printf("records discarded (--out /dev/null): no sidecar written\n");
  } else {
    write_metadata(a, wall, rate, n_failed);
    std::printf("wrote %s and %s.meta.json\n", a.out.c_str(), a.out.c_str());
  }
  if (hs) ce_host_sampler_destroy(hs);
  ce_engine_destroy(eng);
  return EXIT_SUCCESS;
}
