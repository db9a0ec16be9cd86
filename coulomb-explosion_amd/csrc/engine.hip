// Host side of the C-ABI (include/coulomb_explosion.h): engine lifetime, device
// buffers, launch shape, and the per-(precision, N) kernel dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "coulomb_explosion.h"
#include "kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define CE_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(CE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

template <class T>
int hip_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  CE_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return CE_OK;
}

}  // namespace

namespace ce {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace ce

struct ce_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t k0 = nullptr, k1 = nullptr, s0 = nullptr, s1 = nullptr;
  int n = 0;
  double masses[CE_MAX_ATOMS] = {};
  double charges[CE_MAX_ATOMS] = {};
  double coulomb_k = 1.0;
  double rtol = 0, atol = 0;
  size_t capacity = 0;
  size_t n_loaded = 0;
  bool sampled = false;
  // device buffers
  double* geom = nullptr;
  uint8_t* gstat = nullptr;
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
  uint8_t* records = nullptr;   // device staging of packed records (lazy)
  size_t records_cap = 0;       // sims it holds
  hipEvent_t p0 = nullptr, p1 = nullptr;
  // launch
  ce_launch_config launch = {0, 0, 0, 0};
  int n_cus = 0;
  ce_run_stats stats = {};
  bool have_results = false;
};

namespace {

template <class T, int N>
ce::MolConst<T, N> make_const(const ce_engine* e) {
  ce::MolConst<T, N> mc{};
  for (int i = 0; i < N; ++i) {
    mc.im[i] = static_cast<T>(1.0 / e->masses[i]);
    mc.md[i] = e->masses[i];
    for (int j = i + 1; j < N; ++j) {
      const int p = ce::pair_index<N>(i, j);
      mc.pcd[p] = e->coulomb_k * e->charges[i] * e->charges[j];
      mc.pc[p] = static_cast<T>(mc.pcd[p]);
    }
  }
  return mc;
}

template <class T, int N>
int launch_dp5(ce_engine* e, const ce::RunArgs& a, int* grid_out, int* block_out) {
  const auto kern = ce::dp5_persistent<T, N>;
  const int block = e->launch.block_threads > 0 ? e->launch.block_threads : 256;
  int per_cu = e->launch.blocks_per_cu;
  if (per_cu <= 0) {
    int occ = 0;
    CE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, 0));
    per_cu = std::max(occ, 1);
  }
  const int grid = per_cu * e->n_cus;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, e->stream, a, make_const<T, N>(e));
  CE_HIP(hipGetLastError());
  *grid_out = grid;
  *block_out = block;
  return CE_OK;
}

template <class T>
int dispatch_dp5(ce_engine* e, const ce::RunArgs& a, int* g, int* b) {
  switch (e->n) {
    case 1: return launch_dp5<T, 1>(e, a, g, b);
    case 2: return launch_dp5<T, 2>(e, a, g, b);
    case 3: return launch_dp5<T, 3>(e, a, g, b);
    case 4: return launch_dp5<T, 4>(e, a, g, b);
    case 5: return launch_dp5<T, 5>(e, a, g, b);
    case 6: return launch_dp5<T, 6>(e, a, g, b);
    case 7: return launch_dp5<T, 7>(e, a, g, b);
    case 8: return launch_dp5<T, 8>(e, a, g, b);
    case 9: return launch_dp5<T, 9>(e, a, g, b);
    case 10: return launch_dp5<T, 10>(e, a, g, b);
    default: return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
  }
}

template <int N>
void launch_sample(const ce::SampleArgs& a, hipStream_t s) {
  const int block = 256;
  const uint64_t blocks = (a.n_sims + block - 1) / block;
  const int grid = static_cast<int>(std::min<uint64_t>(blocks, 65536));
  hipLaunchKernelGGL(ce::sample_sphere<N>, dim3(grid), dim3(block), 0, s, a);
}

int check_config(const ce_run_config* c) {
  if (c == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "run config is NULL");
  if (c->stop_rule != CE_STOP_PE_FRACTION && c->stop_rule != CE_STOP_MOMENTUM_OR_TMAX)
    return fail(CE_ERR_INVALID_ARGUMENT, "unknown stop_rule");
  if (c->h0_mode != CE_H0_HAIRER && c->h0_mode != CE_H0_FIXED)
    return fail(CE_ERR_INVALID_ARGUMENT, "unknown h0_mode");
  if (c->precision != CE_PRECISION_F32 && c->precision != CE_PRECISION_F64)
    return fail(CE_ERR_INVALID_ARGUMENT, "unknown precision");
  if (!(c->max_dt > 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "max_dt must be positive");
  if (c->h0_mode == CE_H0_FIXED && !(c->h0 > 0.0))
    return fail(CE_ERR_INVALID_ARGUMENT, "fixed h0 must be positive");
  if (c->stop_rule == CE_STOP_MOMENTUM_OR_TMAX && !(c->t_max > 0.0))
    return fail(CE_ERR_INVALID_ARGUMENT, "t_max must be positive");
  if (c->max_rejections < 0) return fail(CE_ERR_INVALID_ARGUMENT, "max_rejections < 0");
  return CE_OK;
}

int copy_results(ce_engine* e, ce_results* out) {
  if (out == nullptr) return CE_OK;
  if (!e->have_results) return fail(CE_ERR_RUNTIME, "no results: run the engine first");
  const size_t ns = e->n_loaded, n3 = 3 * static_cast<size_t>(e->n);
  auto d2h = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    return dst ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, e->stream) : hipSuccess;
  };
  CE_HIP(d2h(out->momenta, e->mom, ns * n3 * sizeof(double)));
  CE_HIP(d2h(out->momenta_f32, e->mom32, ns * n3 * sizeof(float)));
  CE_HIP(d2h(out->steps, e->steps, ns * sizeof(uint32_t)));
  CE_HIP(d2h(out->attempts, e->attempts, ns * sizeof(uint32_t)));
  CE_HIP(d2h(out->status, e->status, ns));
  CE_HIP(d2h(out->t_final, e->tfin, ns * sizeof(double)));
  CE_HIP(d2h(out->energy_err, e->eerr, ns * sizeof(double)));
  if (out->trace_h != nullptr) {
    if (out->trace_cap != e->trace_cap || e->trace == nullptr)
      return fail(CE_ERR_INVALID_ARGUMENT, "trace_cap does not match the enabled trace");
    CE_HIP(d2h(out->trace_h, e->trace, ns * e->trace_cap * sizeof(double)));
  }
  CE_HIP(hipStreamSynchronize(e->stream));
  return CE_OK;
}

void free_engine(ce_engine* e) {
  if (e == nullptr) return;
  hipSetDevice(e->device);
  void* bufs[] = {e->geom,     e->gstat,  e->ctr,  e->mom,  e->mom32,  e->steps,
                  e->attempts, e->status, e->tfin, e->eerr, e->trace, e->records};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (e->k0) hipEventDestroy(e->k0);
  if (e->k1) hipEventDestroy(e->k1);
  if (e->s0) hipEventDestroy(e->s0);
  if (e->s1) hipEventDestroy(e->s1);
  if (e->p0) hipEventDestroy(e->p0);
  if (e->p1) hipEventDestroy(e->p1);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

}  // namespace

extern "C" {

const char* ce_last_error(void) { return g_err.c_str(); }

int ce_device_count(int* count) {
  if (count == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "count is NULL");
  *count = 0;
  const hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return fail(CE_ERR_NO_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  return CE_OK;
}

void ce_run_config_default(ce_run_config* c) {
  if (c == nullptr) return;
  std::memset(c, 0, sizeof(*c));
  c->pe_stop_fraction = 1e-9;
  c->max_dt = 1e12;
  c->t_max = 1e4;
  c->momentum_tol = 1e-7;
  c->h0 = 1e-3;
  c->max_steps = 1000000;
  c->stop_rule = CE_STOP_PE_FRACTION;
  c->redistribute_energy = 1;
  c->h0_mode = CE_H0_HAIRER;
  c->max_rejections = 100;
  c->precision = CE_PRECISION_F32;
}

int ce_engine_create(const ce_molecule_desc* mol, const ce_tolerances* tol, size_t capacity,
                     ce_engine** out) {
  if (out == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (mol == nullptr || tol == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (mol->n_atoms < 1 || mol->n_atoms > CE_MAX_ATOMS)
    return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms must be in [1, CE_MAX_ATOMS]");
  if (mol->masses == nullptr || mol->charges == nullptr)
    return fail(CE_ERR_INVALID_ARGUMENT, "masses/charges NULL");
  for (int i = 0; i < mol->n_atoms; ++i)
    if (!(mol->masses[i] > 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "masses must be positive");
  if (!(tol->rtol > 0.0) || !(tol->atol >= 0.0))
    return fail(CE_ERR_INVALID_ARGUMENT, "rtol must be > 0 and atol >= 0");
  if (capacity == 0) return fail(CE_ERR_INVALID_ARGUMENT, "capacity must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(CE_ERR_NO_DEVICE, "no HIP device");

  ce_engine* e = new ce_engine();
  CE_HIP(hipGetDevice(&e->device));
  e->n = mol->n_atoms;
  for (int i = 0; i < e->n; ++i) {
    e->masses[i] = mol->masses[i];
    e->charges[i] = mol->charges[i];
  }
  e->coulomb_k = mol->coulomb_k;
  e->rtol = tol->rtol;
  e->atol = tol->atol;
  e->capacity = capacity;
  hipDeviceProp_t prop;
  int rc = CE_OK;
  auto chk = [&](hipError_t h, const char* what) {
    if (h != hipSuccess && rc == CE_OK)
      rc = fail(CE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(h));
  };
  chk(hipGetDeviceProperties(&prop, e->device), "hipGetDeviceProperties");
  if (rc == CE_OK) e->n_cus = prop.multiProcessorCount;
  e->stats.n_cus = e->n_cus;
  chk(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate");
  chk(hipEventCreate(&e->k0), "hipEventCreate");
  chk(hipEventCreate(&e->k1), "hipEventCreate");
  chk(hipEventCreate(&e->s0), "hipEventCreate");
  chk(hipEventCreate(&e->s1), "hipEventCreate");
  chk(hipEventCreate(&e->p0), "hipEventCreate");
  chk(hipEventCreate(&e->p1), "hipEventCreate");
  const size_t n3 = 3 * static_cast<size_t>(e->n);
  if (rc == CE_OK) rc = hip_alloc(&e->geom, capacity * n3);
  if (rc == CE_OK) rc = hip_alloc(&e->gstat, capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->ctr, 8);
  if (rc == CE_OK) rc = hip_alloc(&e->mom, capacity * n3);
  if (rc == CE_OK) rc = hip_alloc(&e->mom32, capacity * n3);
  if (rc == CE_OK) rc = hip_alloc(&e->steps, capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->attempts, capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->status, capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->tfin, capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->eerr, capacity);
  if (rc != CE_OK) {
    const std::string msg = g_err;
    free_engine(e);
    g_err = msg;
    return rc;
  }
  e->launch.refill_threshold = 4;
  e->launch.chunk = 64;
  *out = e;
  return CE_OK;
}

int ce_engine_destroy(ce_engine* e) {
  free_engine(e);
  return CE_OK;
}

int ce_engine_set_geometries(ce_engine* e, const double* pos, size_t n_sims) {
  if (e == nullptr || (pos == nullptr && n_sims > 0))
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_sims > e->capacity) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims exceeds capacity");
  CE_HIP(hipSetDevice(e->device));
  const size_t bytes = n_sims * 3 * static_cast<size_t>(e->n) * sizeof(double);
  if (bytes) CE_HIP(hipMemcpyAsync(e->geom, pos, bytes, hipMemcpyHostToDevice, e->stream));
  CE_HIP(hipStreamSynchronize(e->stream));
  e->n_loaded = n_sims;
  e->sampled = false;
  e->have_results = false;
  return CE_OK;
}

int ce_engine_sample(ce_engine* e, uint64_t seed, uint64_t first_index, size_t n_sims,
                     double radius, double min_sep, uint32_t max_attempts) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (!(radius > 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "radius must be positive");
  if (!(min_sep >= 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "min_separation must be >= 0");
  if (max_attempts == 0) return fail(CE_ERR_INVALID_ARGUMENT, "max_attempts must be positive");
  if (n_sims > e->capacity) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims exceeds capacity");
  CE_HIP(hipSetDevice(e->device));
  ce::SampleArgs a{};
  a.geom = e->geom;
  a.status = e->gstat;
  a.first_index = first_index;
  a.n_sims = n_sims;
  a.key0 = static_cast<uint32_t>(seed);
  a.key1 = static_cast<uint32_t>(seed >> 32);
  a.R = static_cast<float>(radius);
  a.twoR = static_cast<float>(2.0 * radius);
  a.R2 = a.R * a.R;
  const float ms = static_cast<float>(min_sep);
  a.min2 = ms * ms;
  a.max_attempts = max_attempts;
  a.n_atoms = e->n;
  CE_HIP(hipEventRecord(e->s0, e->stream));
  if (n_sims > 0) {
    switch (e->n) {
      case 1: launch_sample<1>(a, e->stream); break;
      case 2: launch_sample<2>(a, e->stream); break;
      case 3: launch_sample<3>(a, e->stream); break;
      case 4: launch_sample<4>(a, e->stream); break;
      case 5: launch_sample<5>(a, e->stream); break;
      case 6: launch_sample<6>(a, e->stream); break;
      case 7: launch_sample<7>(a, e->stream); break;
      case 8: launch_sample<8>(a, e->stream); break;
      case 9: launch_sample<9>(a, e->stream); break;
      case 10: launch_sample<10>(a, e->stream); break;
      default: return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
    }
    CE_HIP(hipGetLastError());
  }
  CE_HIP(hipEventRecord(e->s1, e->stream));
  CE_HIP(hipStreamSynchronize(e->stream));
  float ms_elapsed = 0.0f;
  CE_HIP(hipEventElapsedTime(&ms_elapsed, e->s0, e->s1));
  e->stats.sampler_ms = ms_elapsed;
  e->n_loaded = n_sims;
  e->sampled = true;
  e->have_results = false;
  return CE_OK;
}

int ce_engine_get_geometries(ce_engine* e, double* pos, size_t n_sims) {
  if (e == nullptr || pos == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_sims > e->n_loaded) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims exceeds loaded sims");
  CE_HIP(hipSetDevice(e->device));
  CE_HIP(hipMemcpyAsync(pos, e->geom, n_sims * 3 * static_cast<size_t>(e->n) * sizeof(double),
                        hipMemcpyDeviceToHost, e->stream));
  CE_HIP(hipStreamSynchronize(e->stream));
  return CE_OK;
}

int ce_engine_enable_trace(ce_engine* e, uint32_t cap) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  CE_HIP(hipSetDevice(e->device));
  if (e->trace) CE_HIP(hipFree(e->trace));
  e->trace = nullptr;
  e->trace_cap = 0;
  if (cap > 0) {
    const int rc = hip_alloc(&e->trace, e->capacity * cap);
    if (rc != CE_OK) return rc;
    e->trace_cap = cap;
  }
  return CE_OK;
}

int ce_engine_set_launch(ce_engine* e, const ce_launch_config* lc) {
  if (e == nullptr || lc == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (lc->block_threads != 0 &&
      (lc->block_threads % 64 != 0 || lc->block_threads < 64 || lc->block_threads > 256))
    return fail(CE_ERR_INVALID_ARGUMENT, "block_threads must be 64, 128, 192 or 256");
  if (lc->blocks_per_cu < 0 || lc->blocks_per_cu > 64)
    return fail(CE_ERR_INVALID_ARGUMENT, "blocks_per_cu out of range");
  if (lc->refill_threshold < 0 || lc->refill_threshold > 64)
    return fail(CE_ERR_INVALID_ARGUMENT, "refill_threshold must be in [1, 64]");
  if (lc->chunk != 0 && (lc->chunk < 64 || lc->chunk > (1 << 20)))
    return fail(CE_ERR_INVALID_ARGUMENT, "chunk must be in [64, 2^20]");
  e->launch = *lc;
  if (e->launch.refill_threshold == 0) e->launch.refill_threshold = 4;
  if (e->launch.chunk == 0) e->launch.chunk = 64;
  return CE_OK;
}

int ce_engine_get_launch(const ce_engine* e, ce_launch_config* lc) {
  if (e == nullptr || lc == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  *lc = e->launch;
  return CE_OK;
}

int ce_engine_run(ce_engine* e, const ce_run_config* cfg, ce_results* out) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  int rc = check_config(cfg);
  if (rc != CE_OK) return rc;
  CE_HIP(hipSetDevice(e->device));
  ce::RunArgs a{};
  a.geom = e->geom;
  a.geom_status = e->sampled ? e->gstat : nullptr;
  a.n_sims = e->n_loaded;
  a.ctr = e->ctr;
  a.mom = e->mom;
  a.mom32 = e->mom32;
  a.steps = e->steps;
  a.attempts = e->attempts;
  a.status = e->status;
  a.t_final = e->tfin;
  a.eerr = e->eerr;
  a.trace_h = e->trace;
  a.trace_cap = e->trace_cap;
  a.stop_rule = cfg->stop_rule;
  a.rtol = e->rtol;
  a.atol = e->atol;
  a.max_dt = cfg->max_dt;
  a.t_max = cfg->t_max;
  a.pe_stop_fraction = cfg->pe_stop_fraction;
  a.momentum_tol = cfg->momentum_tol;
  a.h0 = cfg->h0;
  a.max_steps = cfg->max_steps;
  a.redistribute = cfg->redistribute_energy ? 1 : 0;
  a.h0_mode = cfg->h0_mode;
  a.max_rej = cfg->max_rejections;
  a.refill_threshold = std::max(1, std::min(64, e->launch.refill_threshold));
  a.chunk = std::max(64, e->launch.chunk);
  // Clear the queue head and totals with a kernel (no memset node, so the
  // sequence stays valid inside a captured graph).
  hipLaunchKernelGGL(ce::zero_u64, dim3(1), dim3(64), 0, e->stream, e->ctr, 8);
  CE_HIP(hipGetLastError());
  CE_HIP(hipEventRecord(e->k0, e->stream));
  int grid = 0, block = 0;
  if (e->n_loaded > 0) {
    rc = (cfg->precision == CE_PRECISION_F64) ? dispatch_dp5<double>(e, a, &grid, &block)
                                              : dispatch_dp5<float>(e, a, &grid, &block);
    if (rc != CE_OK) return rc;
  }
  CE_HIP(hipEventRecord(e->k1, e->stream));
  unsigned long long host_ctr[8] = {};
  CE_HIP(hipMemcpyAsync(host_ctr, e->ctr, sizeof(host_ctr), hipMemcpyDeviceToHost, e->stream));
  CE_HIP(hipStreamSynchronize(e->stream));
  float ms = 0.0f;
  CE_HIP(hipEventElapsedTime(&ms, e->k0, e->k1));
  e->stats.kernel_ms = ms;
  e->stats.sims = e->n_loaded;
  e->stats.attempted_steps = host_ctr[1];
  e->stats.accepted_steps = host_ctr[2];
  e->stats.refills = host_ctr[3];
  e->stats.grid_blocks = grid;
  e->stats.block_threads = block;
  e->stats.n_cus = e->n_cus;
  e->have_results = true;
  return copy_results(e, out);
}

int ce_engine_download(ce_engine* e, ce_results* out) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  CE_HIP(hipSetDevice(e->device));
  return copy_results(e, out);
}

int ce_engine_stats(const ce_engine* e, ce_run_stats* out) {
  if (e == nullptr || out == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = e->stats;
  return CE_OK;
}

size_t ce_record_bytes(int32_t n_atoms) {
  return n_atoms > 0 ? static_cast<size_t>(24 * n_atoms + 5) : 0;
}

int ce_engine_pack_records(ce_engine* e, uint64_t first, size_t n_sims, uint8_t* host_out) {
  if (e == nullptr || (host_out == nullptr && n_sims > 0))
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!e->have_results) return fail(CE_ERR_RUNTIME, "no results: run the engine first");
  if (first > e->n_loaded || n_sims > e->n_loaded - first)
    return fail(CE_ERR_INVALID_ARGUMENT, "record range exceeds the loaded sims");
  if (n_sims == 0) return CE_OK;
  CE_HIP(hipSetDevice(e->device));
  const size_t rb = ce_record_bytes(e->n);
  if (e->records == nullptr || e->records_cap < n_sims) {
    if (e->records) CE_HIP(hipFree(e->records));
    e->records = nullptr;
    e->records_cap = 0;
    const int rc = hip_alloc(&e->records, n_sims * rb);
    if (rc != CE_OK) return rc;
    e->records_cap = n_sims;
  }
  const dim3 grid(static_cast<unsigned>((n_sims + ce::kRecBlock - 1) / ce::kRecBlock));
  const dim3 block(ce::kRecBlock);
  CE_HIP(hipEventRecord(e->p0, e->stream));
  switch (e->n) {
#define CE_PACK(NN)                                                                            \
  case NN:                                                                                     \
    hipLaunchKernelGGL(ce::pack_records<NN>, grid, block, 0, e->stream, e->geom, e->mom32,     \
                       e->steps, e->status, first, static_cast<uint64_t>(n_sims), e->records); \
    break;
    CE_PACK(1) CE_PACK(2) CE_PACK(3) CE_PACK(4) CE_PACK(5) CE_PACK(6) CE_PACK(7) CE_PACK(8)
    CE_PACK(9) CE_PACK(10)
#undef CE_PACK
    default: return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
  }
  CE_HIP(hipGetLastError());
  CE_HIP(hipEventRecord(e->p1, e->stream));
  CE_HIP(hipMemcpyAsync(host_out, e->records, n_sims * rb, hipMemcpyDeviceToHost, e->stream));
  CE_HIP(hipStreamSynchronize(e->stream));
  float ms = 0.0f;
  CE_HIP(hipEventElapsedTime(&ms, e->p0, e->p1));
  e->stats.pack_ms = ms;
  return CE_OK;
}

int ce_batched_accelerations_f32(size_t n, size_t k, const float* px, const float* py,
                                 const float* pz, const float* pair_const, const float* inv_mass,
                                 float* ax, float* ay, float* az) {
  if (n < 1 || n > CE_MAX_ATOMS) return fail(CE_ERR_INVALID_ARGUMENT, "n out of range");
  if (!px || !py || !pz || !pair_const || !inv_mass || !ax || !ay || !az)
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (k == 0) return CE_OK;
  const size_t nk = n * k;
  float *d = nullptr;
  const size_t total = 6 * nk + n * n + n;
  CE_HIP(hipMalloc(&d, total * sizeof(float)));
  float *dx = d, *dy = d + nk, *dz = d + 2 * nk, *dax = d + 3 * nk, *day = d + 4 * nk,
        *daz = d + 5 * nk, *dpc = d + 6 * nk, *dim = dpc + n * n;
  int rc = CE_OK;
  auto chk = [&](hipError_t h) {
    if (h != hipSuccess && rc == CE_OK) rc = fail(CE_ERR_HIP, hipGetErrorString(h));
  };
  chk(hipMemcpy(dx, px, nk * sizeof(float), hipMemcpyHostToDevice));
  chk(hipMemcpy(dy, py, nk * sizeof(float), hipMemcpyHostToDevice));
  chk(hipMemcpy(dz, pz, nk * sizeof(float), hipMemcpyHostToDevice));
  chk(hipMemcpy(dpc, pair_const, n * n * sizeof(float), hipMemcpyHostToDevice));
  chk(hipMemcpy(dim, inv_mass, n * sizeof(float), hipMemcpyHostToDevice));
  if (rc == CE_OK) {
    const dim3 grid(static_cast<unsigned>((k + 255) / 256)), block(256);
    switch (n) {
#define CE_ACC(NN) \
  case NN: hipLaunchKernelGGL(ce::soa_accel<NN>, grid, block, 0, nullptr, k, dx, dy, dz, dpc, dim, dax, day, daz); break;
      CE_ACC(1) CE_ACC(2) CE_ACC(3) CE_ACC(4) CE_ACC(5) CE_ACC(6) CE_ACC(7) CE_ACC(8) CE_ACC(9)
      CE_ACC(10)
#undef CE_ACC
    }
    chk(hipGetLastError());
    chk(hipMemcpy(ax, dax, nk * sizeof(float), hipMemcpyDeviceToHost));
    chk(hipMemcpy(ay, day, nk * sizeof(float), hipMemcpyDeviceToHost));
    chk(hipMemcpy(az, daz, nk * sizeof(float), hipMemcpyDeviceToHost));
  }
  hipFree(d);
  return rc;
}

int ce_batched_potential_energy_f32(size_t n, size_t k, const float* px, const float* py,
                                    const float* pz, const float* pair_const, float* pe) {
  if (n < 1 || n > CE_MAX_ATOMS) return fail(CE_ERR_INVALID_ARGUMENT, "n out of range");
  if (!px || !py || !pz || !pair_const || !pe) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (k == 0) return CE_OK;
  const size_t nk = n * k;
  float* d = nullptr;
  CE_HIP(hipMalloc(&d, (3 * nk + n * n + k) * sizeof(float)));
  float *dx = d, *dy = d + nk, *dz = d + 2 * nk, *dpc = d + 3 * nk, *dpe = dpc + n * n;
  int rc = CE_OK;
  auto chk = [&](hipError_t h) {
    if (h != hipSuccess && rc == CE_OK) rc = fail(CE_ERR_HIP, hipGetErrorString(h));
  };
  chk(hipMemcpy(dx, px, nk * sizeof(float), hipMemcpyHostToDevice));
  chk(hipMemcpy(dy, py, nk * sizeof(float), hipMemcpyHostToDevice));
  chk(hipMemcpy(dz, pz, nk * sizeof(float), hipMemcpyHostToDevice));
  chk(hipMemcpy(dpc, pair_const, n * n * sizeof(float), hipMemcpyHostToDevice));
  if (rc == CE_OK) {
    const dim3 grid(static_cast<unsigned>((k + 255) / 256)), block(256);
    switch (n) {
#define CE_PE(NN) \
  case NN: hipLaunchKernelGGL(ce::soa_pe<NN>, grid, block, 0, nullptr, k, dx, dy, dz, dpc, dpe); break;
      CE_PE(1) CE_PE(2) CE_PE(3) CE_PE(4) CE_PE(5) CE_PE(6) CE_PE(7) CE_PE(8) CE_PE(9) CE_PE(10)
#undef CE_PE
    }
    chk(hipGetLastError());
    chk(hipMemcpy(pe, dpe, k * sizeof(float), hipMemcpyDeviceToHost));
  }
  hipFree(d);
  return rc;
}

}  // extern "C"
