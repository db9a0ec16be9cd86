// Host side of the C-ABI (include/coulomb_explosion.h): engine lifetime, device
// buffers, species / segment layout, launch shape, and the kernel dispatch.
//
// An engine holds 1..4 species (molecule types).  Species s owns the global sim
// range [start_s, start_s + count_s); per-sim arrays (steps, status, ...) are
// indexed by global sim, per-atom arrays (geometry, momenta) are species-major:
// species s starts at pos_off_s = sum_{r<s} count_r * 3 n_r.  A single-molecule
// engine (ce_engine_create) is one species whose count is the loaded batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "coulomb_explosion.h"
#include "dispatch.hpp"
#include "engine.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define CE_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ce::clear_stale_error(); /* reported here: not left for the next check */   \
      return fail(CE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                             \
  } while (0)

template <class T>
int hip_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (count > SIZE_MAX / sizeof(T))
    return fail(CE_ERR_INVALID_ARGUMENT, "buffer size overflows size_t");
  CE_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return CE_OK;
}

// a * b + c without wrapping (false on overflow).
bool mul_add_ok(size_t a, size_t b, size_t c, size_t* out) {
  size_t m = 0;
  if (__builtin_mul_overflow(a, b, &m)) return false;
  return !__builtin_add_overflow(m, c, out);
}

}  // namespace

namespace ce {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace ce


namespace {

void free_engine(ce_engine* e) {
  if (e == nullptr) return;
  (void)hipSetDevice(e->device);
  void* bufs[] = {e->geom,   e->gstat,    e->qsim,   e->ctr,  e->mom,     e->mom32,
                  e->steps,  e->attempts, e->status, e->tfin, e->eerr,    e->trace,
                  e->records, e->state_x, e->state_v, e->state_h, e->state_e0, e->list};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (e->copy_stream) (void)hipStreamSynchronize(e->copy_stream);
  for (int s = 0; s < 2; ++s) {
    if (e->rec_async[s]) (void)hipFree(e->rec_async[s]);
    if (e->rec_packed[s]) (void)hipEventDestroy(e->rec_packed[s]);
    if (e->rec_copied[s]) (void)hipEventDestroy(e->rec_copied[s]);
  }
  if (e->copy_stream) (void)hipStreamDestroy(e->copy_stream);
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->first_dev) (void)hipFree(e->first_dev);
  hipEvent_t evs[] = {e->k0, e->k1, e->s0, e->s1, e->p0, e->p1};
  for (hipEvent_t v : evs)
    if (v) (void)hipEventDestroy(v);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int check_species(const double* masses, int n) {
  if (n < 1 || n > CE_MAX_ATOMS) return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms must be in [1, CE_MAX_ATOMS]");
  for (int i = 0; i < n; ++i)
    if (!(masses[i] > 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "masses must be positive");
  return CE_OK;
}

// Allocates the engine's device state for its species (capacities set).
int init_engine(ce_engine* e, const ce_tolerances* tol) {
  if (!(tol->rtol > 0.0) || !(tol->atol >= 0.0))
    return fail(CE_ERR_INVALID_ARGUMENT, "rtol must be > 0 and atol >= 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(CE_ERR_NO_DEVICE, "no HIP device");
  CE_HIP(hipGetDevice(&e->device));
  e->rtol = tol->rtol;
  e->atol = tol->atol;
  e->capacity = 0;
  e->total_pos = 0;
  e->total_q = 0;
  for (int s = 0; s < e->n_species; ++s) {
    const size_t cap = e->sp[s].capacity, n = static_cast<size_t>(e->sp[s].n);
    if (!mul_add_ok(cap, 1, e->capacity, &e->capacity) ||
        !mul_add_ok(cap, 3 * n, e->total_pos, &e->total_pos) ||
        !mul_add_ok(cap, n, e->total_q, &e->total_q) || e->total_pos > SIZE_MAX / sizeof(double))
      return fail(CE_ERR_INVALID_ARGUMENT, "capacity too large (buffer sizes overflow)");
  }
  hipDeviceProp_t prop;
  CE_HIP(hipGetDeviceProperties(&prop, e->device));
  e->n_cus = prop.multiProcessorCount;
  e->stats.n_cus = e->n_cus;
  CE_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  hipEvent_t* evs[] = {&e->k0, &e->k1, &e->s0, &e->s1, &e->p0, &e->p1};
  for (hipEvent_t* v : evs) CE_HIP(hipEventCreate(v));
  int rc = CE_OK;
  if (rc == CE_OK) rc = hip_alloc(&e->geom, e->total_pos);
  if (rc == CE_OK) rc = hip_alloc(&e->gstat, e->capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->ctr, ce::kCtrWords);
  if (rc == CE_OK) rc = hip_alloc(&e->mom, e->total_pos);
  if (rc == CE_OK) rc = hip_alloc(&e->mom32, e->total_pos);
  if (rc == CE_OK) rc = hip_alloc(&e->steps, e->capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->attempts, e->capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->status, e->capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->tfin, e->capacity);
  if (rc == CE_OK) rc = hip_alloc(&e->eerr, e->capacity);
  bool any_levels = false;
  for (int s = 0; s < e->n_species; ++s) any_levels = any_levels || e->sp[s].levels > 0;
  if (rc == CE_OK && any_levels) rc = hip_alloc(&e->qsim, e->total_q);
  return rc;
}

int finish_create(ce_engine* e, int rc, ce_engine** out) {
  if (rc != CE_OK) {
    const std::string msg = g_err;
    free_engine(e);
    g_err = msg;
    return rc;
  }
  *out = e;
  return CE_OK;
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
  if (c->stop_rule == CE_STOP_MOMENTUM_OR_TMAX && !(c->momentum_tol >= 0.0))
    return fail(CE_ERR_INVALID_ARGUMENT, "momentum_tol must be >= 0");
  if (c->max_rejections < 0) return fail(CE_ERR_INVALID_ARGUMENT, "max_rejections < 0");
  return CE_OK;
}

// Species-major per-atom arrays are packed at the *capacity* offsets on the
// device; the host sees them packed at the *loaded* counts.  For a
// single-species engine the two coincide.
int copy_per_atom(ce_engine* e, void* host, const void* dev, size_t elem, int per_atom_width,
                  hipMemcpyKind kind) {
  size_t host_off = 0;
  for (int s = 0; s < e->n_species; ++s) {
    uint64_t start, pos_off, q_off;
    e->offsets(s, &start, &pos_off, &q_off);
    const size_t dev_off = per_atom_width == 3 ? pos_off : q_off;
    const size_t cnt = e->sp[s].count * per_atom_width * e->sp[s].n;
    if (cnt == 0) continue;
    char* h = static_cast<char*>(host) + host_off * elem;
    const char* d = static_cast<const char*>(dev) + dev_off * elem;
    if (kind == hipMemcpyDeviceToHost)
      CE_HIP(hipMemcpyAsync(h, d, cnt * elem, kind, e->stream));
    else
      CE_HIP(hipMemcpyAsync(const_cast<char*>(d), h, cnt * elem, kind, e->stream));
    host_off += cnt;
  }
  return CE_OK;
}

// Per-sim arrays: the device keeps species s at global index start_s (capacity
// based); the host sees the loaded sims back to back.
int copy_per_sim(ce_engine* e, void* host, const void* dev, size_t elem, hipMemcpyKind kind) {
  size_t host_off = 0;
  for (int s = 0; s < e->n_species; ++s) {
    uint64_t start, pos_off, q_off;
    e->offsets(s, &start, &pos_off, &q_off);
    const size_t cnt = e->sp[s].count;
    if (cnt == 0) continue;
    char* h = static_cast<char*>(host) + host_off * elem;
    const char* d = static_cast<const char*>(dev) + start * elem;
    if (kind == hipMemcpyDeviceToHost)
      CE_HIP(hipMemcpyAsync(h, d, cnt * elem, kind, e->stream));
    else
      CE_HIP(hipMemcpyAsync(const_cast<char*>(d), h, cnt * elem, kind, e->stream));
    host_off += cnt;
  }
  return CE_OK;
}

int copy_results(ce_engine* e, ce_results* out) {
  if (out == nullptr) return CE_OK;
  if (!e->have_results) return fail(CE_ERR_RUNTIME, "no results: run the engine first");
  // Checked before any copy is queued, so an error leaves no copy in flight
  // into the caller's buffers.
  if (out->energy_err && !e->eerr_valid)
    return fail(CE_ERR_INVALID_ARGUMENT,
                "energy_err: the last run did not compute it (ce_engine_enable_energy_diagnostic)");
  if (out->trace_h != nullptr && (out->trace_cap != e->trace_cap || e->trace == nullptr))
    return fail(CE_ERR_INVALID_ARGUMENT, "trace_cap does not match the enabled trace");
  int rc = CE_OK;
  const hipMemcpyKind d2h = hipMemcpyDeviceToHost;
  if (rc == CE_OK && out->momenta) rc = copy_per_atom(e, out->momenta, e->mom, sizeof(double), 3, d2h);
  if (rc == CE_OK && out->momenta_f32) rc = copy_per_atom(e, out->momenta_f32, e->mom32, sizeof(float), 3, d2h);
  if (rc == CE_OK && out->steps) rc = copy_per_sim(e, out->steps, e->steps, sizeof(uint32_t), d2h);
  if (rc == CE_OK && out->attempts) rc = copy_per_sim(e, out->attempts, e->attempts, sizeof(uint32_t), d2h);
  if (rc == CE_OK && out->status) rc = copy_per_sim(e, out->status, e->status, 1, d2h);
  if (rc == CE_OK && out->t_final) rc = copy_per_sim(e, out->t_final, e->tfin, sizeof(double), d2h);
  if (rc == CE_OK && out->energy_err)
    rc = copy_per_sim(e, out->energy_err, e->eerr, sizeof(double), d2h);
  if (rc == CE_OK && out->trace_h != nullptr)
    rc = copy_per_sim(e, out->trace_h, e->trace, sizeof(double) * e->trace_cap, d2h);
  // Wait for whatever was queued, also after a failed enqueue.
  const hipError_t se = hipStreamSynchronize(e->stream);
  if (rc != CE_OK) return rc;
  if (se != hipSuccess) return fail(CE_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
  return CE_OK;
}

// The reference-layout force and potential-energy kernels over host buffers
// (one allocation per call; not on the engine's path).
template <class T>
int soa_accelerations(size_t n, size_t k, const T* px, const T* py, const T* pz,
                      const T* pair_const, const T* inv_mass, T* ax, T* ay, T* az) {
  if (n < 1 || n > CE_MAX_ATOMS) return fail(CE_ERR_INVALID_ARGUMENT, "n out of range");
  if (!px || !py || !pz || !pair_const || !inv_mass || !ax || !ay || !az)
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (k == 0) return CE_OK;
  if (k > SIZE_MAX / (8 * sizeof(T) * CE_MAX_ATOMS))
    return fail(CE_ERR_INVALID_ARGUMENT, "k too large");
  const size_t nk = n * k;
  T* d = nullptr;
  const size_t total = 6 * nk + n * n + n;
  CE_HIP(hipMalloc(&d, total * sizeof(T)));
  T *dx = d, *dy = d + nk, *dz = d + 2 * nk, *dax = d + 3 * nk, *day = d + 4 * nk,
    *daz = d + 5 * nk, *dpc = d + 6 * nk, *dim = dpc + n * n;
  int rc = CE_OK;
  auto chk = [&](hipError_t h) {
    if (h == hipSuccess) return;
    ce::clear_stale_error();
    if (rc == CE_OK) rc = fail(CE_ERR_HIP, hipGetErrorString(h));
  };
  chk(hipMemcpy(dx, px, nk * sizeof(T), hipMemcpyHostToDevice));
  chk(hipMemcpy(dy, py, nk * sizeof(T), hipMemcpyHostToDevice));
  chk(hipMemcpy(dz, pz, nk * sizeof(T), hipMemcpyHostToDevice));
  chk(hipMemcpy(dpc, pair_const, n * n * sizeof(T), hipMemcpyHostToDevice));
  chk(hipMemcpy(dim, inv_mass, n * sizeof(T), hipMemcpyHostToDevice));
  if (rc == CE_OK) {
    const dim3 grid(static_cast<unsigned>((k + 255) / 256)), block(256);
    ce::clear_stale_error();
    switch (n) {
#define CE_ACC(NN)                                                                           \
  case NN:                                                                                   \
    hipLaunchKernelGGL((ce::soa_accel<T, NN>), grid, block, 0, nullptr, k, dx, dy, dz, dpc,  \
                       dim, dax, day, daz);                                                  \
    break;
      CE_ACC(1) CE_ACC(2) CE_ACC(3) CE_ACC(4) CE_ACC(5) CE_ACC(6) CE_ACC(7) CE_ACC(8) CE_ACC(9)
      CE_ACC(10)
#undef CE_ACC
    }
    chk(hipGetLastError());
    chk(hipMemcpy(ax, dax, nk * sizeof(T), hipMemcpyDeviceToHost));
    chk(hipMemcpy(ay, day, nk * sizeof(T), hipMemcpyDeviceToHost));
    chk(hipMemcpy(az, daz, nk * sizeof(T), hipMemcpyDeviceToHost));
  }
  (void)hipFree(d);
  return rc;
}

template <class T>
int soa_potential_energy(size_t n, size_t k, const T* px, const T* py, const T* pz,
                         const T* pair_const, T* pe) {
  if (n < 1 || n > CE_MAX_ATOMS) return fail(CE_ERR_INVALID_ARGUMENT, "n out of range");
  if (!px || !py || !pz || !pair_const || !pe) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (k == 0) return CE_OK;
  if (k > SIZE_MAX / (8 * sizeof(T) * CE_MAX_ATOMS))
    return fail(CE_ERR_INVALID_ARGUMENT, "k too large");
  const size_t nk = n * k;
  T* d = nullptr;
  CE_HIP(hipMalloc(&d, (3 * nk + n * n + k) * sizeof(T)));
  T *dx = d, *dy = d + nk, *dz = d + 2 * nk, *dpc = d + 3 * nk, *dpe = dpc + n * n;
  int rc = CE_OK;
  auto chk = [&](hipError_t h) {
    if (h == hipSuccess) return;
    ce::clear_stale_error();
    if (rc == CE_OK) rc = fail(CE_ERR_HIP, hipGetErrorString(h));
  };
  chk(hipMemcpy(dx, px, nk * sizeof(T), hipMemcpyHostToDevice));
  chk(hipMemcpy(dy, py, nk * sizeof(T), hipMemcpyHostToDevice));
  chk(hipMemcpy(dz, pz, nk * sizeof(T), hipMemcpyHostToDevice));
  chk(hipMemcpy(dpc, pair_const, n * n * sizeof(T), hipMemcpyHostToDevice));
  if (rc == CE_OK) {
    const dim3 grid(static_cast<unsigned>((k + 255) / 256)), block(256);
    ce::clear_stale_error();
    switch (n) {
#define CE_PE(NN)                                                                               \
  case NN:                                                                                      \
    hipLaunchKernelGGL((ce::soa_pe<T, NN>), grid, block, 0, nullptr, k, dx, dy, dz, dpc, dpe); \
    break;
      CE_PE(1) CE_PE(2) CE_PE(3) CE_PE(4) CE_PE(5) CE_PE(6) CE_PE(7) CE_PE(8) CE_PE(9) CE_PE(10)
#undef CE_PE
    }
    chk(hipGetLastError());
    chk(hipMemcpy(pe, dpe, k * sizeof(T), hipMemcpyDeviceToHost));
  }
  (void)hipFree(d);
  return rc;
}

template <int N>
void launch_sample(const ce::SampleArgs& a, hipStream_t s) {
  const int block = 256;
  const uint64_t blocks = (a.n_sims + block - 1) / block;
  const int grid = static_cast<int>(std::min<uint64_t>(blocks, 65536));
  hipLaunchKernelGGL(ce::sample_sphere<N>, dim3(grid), dim3(block), 0, s, a);
}

int sample_segment(ce_engine* e, int s, uint64_t seed, uint64_t first_global, double radius,
                   double min_sep, uint32_t max_attempts, bool draw_charges,
                   const uint64_t* first_dev = nullptr) {
  uint64_t start, pos_off, q_off;
  e->offsets(s, &start, &pos_off, &q_off);
  ce::SampleArgs a{};
  a.geom = e->geom + pos_off;
  a.status = e->gstat + start;
  a.charges = draw_charges ? e->qsim + q_off : nullptr;
  a.first_index = first_global;
  a.first_index_dev = first_dev;
  a.n_sims = e->sp[s].count;
  a.key0 = static_cast<uint32_t>(seed);
  a.key1 = static_cast<uint32_t>(seed >> 32);
  a.R = static_cast<float>(radius);
  a.twoR = static_cast<float>(2.0 * radius);
  a.R2 = a.R * a.R;
  const float ms = static_cast<float>(min_sep);
  a.min2 = ms * ms;
  a.max_attempts = max_attempts;
  a.n_atoms = e->sp[s].n;
  a.charge_levels = e->sp[s].levels;
  a.max_total_charge = e->sp[s].max_total;
  if (a.n_sims == 0) return CE_OK;
  ce::clear_stale_error();
  switch (e->sp[s].n) {
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
  return CE_OK;
}

int finish_sampling(ce_engine* e) {
  CE_HIP(hipEventRecord(e->s1, e->stream));
  CE_HIP(hipStreamSynchronize(e->stream));
  float ms = 0.0f;
  CE_HIP(hipEventElapsedTime(&ms, e->s0, e->s1));
  e->stats.sampler_ms = ms;
  e->sampled = true;
  e->have_results = false;
  return CE_OK;
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
  int rc = check_species(mol->masses, mol->n_atoms);
  if (rc != CE_OK) return rc;
  if (!(tol->rtol > 0.0) || !(tol->atol >= 0.0))
    return fail(CE_ERR_INVALID_ARGUMENT, "rtol must be > 0 and atol >= 0");
  if (capacity == 0) return fail(CE_ERR_INVALID_ARGUMENT, "capacity must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(CE_ERR_NO_DEVICE, "no HIP device");
  ce_engine* e = new ce_engine();
  e->n_species = 1;
  ce::Species& s = e->sp[0];
  s.n = mol->n_atoms;
  for (int i = 0; i < s.n; ++i) {
    s.masses[i] = mol->masses[i];
    s.charges[i] = mol->charges[i];
  }
  s.capacity = capacity;
  e->coulomb_k = mol->coulomb_k;
  return finish_create(e, init_engine(e, tol), out);
}

int ce_engine_create_mixed(const ce_species_desc* species, int32_t n_species, double coulomb_k,
                           const ce_tolerances* tol, ce_engine** out) {
  if (out == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (species == nullptr || tol == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_species < 1 || n_species > CE_MAX_SPECIES)
    return fail(CE_ERR_INVALID_ARGUMENT, "n_species must be in [1, CE_MAX_SPECIES]");
  for (int s = 0; s < n_species; ++s) {
    const ce_species_desc& d = species[s];
    if (d.masses == nullptr || (d.charges == nullptr && d.charge_levels == 0))
      return fail(CE_ERR_INVALID_ARGUMENT, "masses/charges NULL");
    const int rc = check_species(d.masses, d.n_atoms);
    if (rc != CE_OK) return rc;
    if (d.charge_levels < 0 || d.charge_levels > 255)
      return fail(CE_ERR_INVALID_ARGUMENT, "charge_levels must be in [0, 255]");
    if (d.count == 0) return fail(CE_ERR_INVALID_ARGUMENT, "species count must be positive");
    if (!(d.radius > 0.0) || !(d.min_separation >= 0.0))
      return fail(CE_ERR_INVALID_ARGUMENT, "species radius must be > 0, min_separation >= 0");
  }
  if (!(tol->rtol > 0.0) || !(tol->atol >= 0.0))
    return fail(CE_ERR_INVALID_ARGUMENT, "rtol must be > 0 and atol >= 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(CE_ERR_NO_DEVICE, "no HIP device");
  ce_engine* e = new ce_engine();
  e->n_species = n_species;
  e->mixed = true;
  e->coulomb_k = coulomb_k;
  for (int s = 0; s < n_species; ++s) {
    const ce_species_desc& d = species[s];
    ce::Species& sp = e->sp[s];
    sp.n = d.n_atoms;
    for (int i = 0; i < sp.n; ++i) {
      sp.masses[i] = d.masses[i];
      sp.charges[i] = d.charges ? d.charges[i] : 1.0;
    }
    sp.radius = d.radius;
    sp.min_sep = d.min_separation;
    sp.levels = d.charge_levels;
    sp.max_total = d.max_total_charge;
    sp.capacity = d.count;
    sp.count = d.count;
  }
  return finish_create(e, init_engine(e, tol), out);
}

int ce_engine_destroy(ce_engine* e) {
  free_engine(e);
  return CE_OK;
}

int ce_engine_set_geometries(ce_engine* e, const double* pos, size_t n_sims) {
  if (e == nullptr || (pos == nullptr && n_sims > 0))
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (e->mixed) {
    if (n_sims != e->capacity)
      return fail(CE_ERR_INVALID_ARGUMENT, "a mixed engine takes all its sims at once");
  } else {
    if (n_sims > e->capacity) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims exceeds capacity");
    e->sp[0].count = n_sims;
  }
  CE_HIP(hipSetDevice(e->device));
  int rc = copy_per_atom(e, const_cast<double*>(pos), e->geom, sizeof(double), 3,
                         hipMemcpyHostToDevice);
  if (rc != CE_OK) return rc;
  CE_HIP(hipStreamSynchronize(e->stream));
  e->sampled = false;
  e->have_charges = false;
  e->have_results = false;
  return CE_OK;
}

int ce_engine_set_charges(ce_engine* e, const double* charges, size_t n_sims) {
  if (e == nullptr || (charges == nullptr && n_sims > 0))
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_sims != e->loaded())
    return fail(CE_ERR_INVALID_ARGUMENT, "charges must cover exactly the loaded sims");
  CE_HIP(hipSetDevice(e->device));
  if (e->qsim == nullptr) {
    const int rc = hip_alloc(&e->qsim, e->total_q);
    if (rc != CE_OK) return rc;
  }
  size_t total = 0;
  for (int s = 0; s < e->n_species; ++s) total += e->sp[s].count * e->sp[s].n;
  std::vector<float> q(total);
  for (size_t i = 0; i < total; ++i) q[i] = static_cast<float>(charges[i]);
  int rc = copy_per_atom(e, q.data(), e->qsim, sizeof(float), 1, hipMemcpyHostToDevice);
  if (rc != CE_OK) return rc;
  CE_HIP(hipStreamSynchronize(e->stream));
  e->have_charges = true;
  e->have_results = false;
  return CE_OK;
}

int ce_engine_get_charges(ce_engine* e, double* charges, size_t n_sims) {
  if (e == nullptr || charges == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!e->have_charges) return fail(CE_ERR_RUNTIME, "the loaded sims carry no per-sim charges");
  if (n_sims != e->loaded())
    return fail(CE_ERR_INVALID_ARGUMENT, "n_sims must equal the loaded sims");
  CE_HIP(hipSetDevice(e->device));
  size_t total = 0;
  for (int s = 0; s < e->n_species; ++s) total += e->sp[s].count * e->sp[s].n;
  std::vector<float> q(total);
  int rc = copy_per_atom(e, q.data(), e->qsim, sizeof(float), 1, hipMemcpyDeviceToHost);
  if (rc != CE_OK) return rc;
  CE_HIP(hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < total; ++i) charges[i] = q[i];
  return CE_OK;
}

int ce_engine_sample(ce_engine* e, uint64_t seed, uint64_t first_index, size_t n_sims,
                     double radius, double min_sep, uint32_t max_attempts) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (e->mixed) return fail(CE_ERR_INVALID_ARGUMENT, "mixed engines sample with ce_engine_sample_species");
  if (!(radius > 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "radius must be positive");
  if (!(min_sep >= 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "min_separation must be >= 0");
  if (max_attempts == 0) return fail(CE_ERR_INVALID_ARGUMENT, "max_attempts must be positive");
  if (n_sims > e->capacity) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims exceeds capacity");
  CE_HIP(hipSetDevice(e->device));
  e->sp[0].count = n_sims;
  CE_HIP(hipEventRecord(e->s0, e->stream));
  const int rc = sample_segment(e, 0, seed, first_index, radius, min_sep, max_attempts, false);
  if (rc != CE_OK) return rc;
  e->have_charges = false;
  return finish_sampling(e);
}

int ce_engine_sample_species(ce_engine* e, uint64_t seed, uint64_t first_index,
                             uint32_t max_attempts) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (!e->mixed) return fail(CE_ERR_INVALID_ARGUMENT, "engine has no species table");
  if (max_attempts == 0) return fail(CE_ERR_INVALID_ARGUMENT, "max_attempts must be positive");
  CE_HIP(hipSetDevice(e->device));
  CE_HIP(hipEventRecord(e->s0, e->stream));
  bool any_levels = false;
  for (int s = 0; s < e->n_species; ++s) {
    uint64_t start, pos_off, q_off;
    e->offsets(s, &start, &pos_off, &q_off);
    const ce::Species& sp = e->sp[s];
    const int rc = sample_segment(e, s, seed, first_index + start, sp.radius, sp.min_sep,
                                  max_attempts, sp.levels > 0);
    if (rc != CE_OK) return rc;
    any_levels = any_levels || sp.levels > 0;
  }
  if (any_levels) {
    // species with fixed charges still need their charges in the per-sim table
    for (int s = 0; s < e->n_species; ++s) {
      const ce::Species& sp = e->sp[s];
      if (sp.levels > 0) continue;
      uint64_t start, pos_off, q_off;
      e->offsets(s, &start, &pos_off, &q_off);
      std::vector<float> q(sp.count * sp.n);
      for (size_t i = 0; i < q.size(); ++i) q[i] = static_cast<float>(sp.charges[i % sp.n]);
      CE_HIP(hipMemcpyAsync(e->qsim + q_off, q.data(), q.size() * sizeof(float),
                            hipMemcpyHostToDevice, e->stream));
      CE_HIP(hipStreamSynchronize(e->stream));
    }
  }
  e->have_charges = any_levels;
  return finish_sampling(e);
}

int ce_engine_get_geometries(ce_engine* e, double* pos, size_t n_sims) {
  if (e == nullptr || pos == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_sims != e->loaded()) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims must equal the loaded sims");
  CE_HIP(hipSetDevice(e->device));
  const int rc = copy_per_atom(e, pos, e->geom, sizeof(double), 3, hipMemcpyDeviceToHost);
  if (rc != CE_OK) return rc;
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

int ce_engine_enable_energy_diagnostic(ce_engine* e, int32_t on) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  e->energy_diag = on != 0;
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
  if (lc->mixed_launch < CE_MIXED_AUTO || lc->mixed_launch > CE_MIXED_PER_SPECIES)
    return fail(CE_ERR_INVALID_ARGUMENT, "mixed_launch must be CE_MIXED_AUTO, _ONE_LAUNCH or _PER_SPECIES");
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
  const int rc = ce::engine_run(e, cfg, nullptr, 0);
  return rc != CE_OK ? rc : copy_results(e, out);
}

int ce_engine_run_subset(ce_engine* e, const ce_run_config* cfg, const uint64_t* sims,
                         size_t n, ce_results* out) {
  if (e == nullptr || (sims == nullptr && n > 0))
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (e->n_species != 1) return fail(CE_ERR_INVALID_ARGUMENT, "subsets need a single-species engine");
  if (n > e->loaded()) return fail(CE_ERR_INVALID_ARGUMENT, "more sims than loaded");
  if (e->loaded() > UINT32_MAX)
    return fail(CE_ERR_INVALID_ARGUMENT, "subsets index at most 2^32 loaded sims");
  std::vector<uint32_t> idx(n);
  for (size_t i = 0; i < n; ++i) {
    if (sims[i] >= e->loaded()) return fail(CE_ERR_INVALID_ARGUMENT, "sim index out of range");
    idx[i] = static_cast<uint32_t>(sims[i]);
  }
  CE_HIP(hipSetDevice(e->device));
  if (e->list == nullptr) {
    const int arc = hip_alloc(&e->list, e->capacity);
    if (arc != CE_OK) return arc;
  }
  if (n > 0) {
    CE_HIP(hipMemcpyAsync(e->list, idx.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice,
                          e->stream));
    CE_HIP(hipStreamSynchronize(e->stream));
  }
  const int rc = ce::engine_run(e, cfg, e->list, n);
  return rc != CE_OK ? rc : copy_results(e, out);
}

int ce_engine_enable_state(ce_engine* e, int32_t enable) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  CE_HIP(hipSetDevice(e->device));
  CE_HIP(hipStreamSynchronize(e->stream));
  if (!enable) {
    void* b[] = {e->state_x, e->state_v, e->state_h, e->state_e0};
    for (void* p : b)
      if (p) CE_HIP(hipFree(p));
    e->state_x = e->state_v = nullptr;
    e->state_h = e->state_e0 = nullptr;
    e->state_prec = -1;
    return CE_OK;
  }
  if (e->state_x != nullptr) return CE_OK;
  double *x = nullptr, *v = nullptr, *h = nullptr, *e0 = nullptr;
  int rc = hip_alloc(&x, e->total_pos);  // room for T = double
  if (rc == CE_OK) rc = hip_alloc(&v, e->total_pos);
  if (rc == CE_OK) rc = hip_alloc(&h, e->capacity);
  if (rc == CE_OK) rc = hip_alloc(&e0, e->capacity);
  if (rc != CE_OK) {
    // all or nothing: a run writes through every state pointer once state_x is set
    for (double* p : {x, v, h, e0})
      if (p) (void)hipFree(p);
    return rc;
  }
  e->state_x = x;
  e->state_v = v;
  e->state_h = h;
  e->state_e0 = e0;
  e->state_prec = -1;
  return CE_OK;
}

int ce_engine_set_state(ce_engine* e, const double* pos, const double* vel, const double* h,
                        const double* e0, size_t n_sims, int32_t precision) {
  if (e == nullptr || ((pos == nullptr || vel == nullptr) && n_sims > 0))
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (precision != CE_PRECISION_F32 && precision != CE_PRECISION_F64)
    return fail(CE_ERR_INVALID_ARGUMENT, "unknown precision");
  if (e->mixed) {
    if (n_sims != e->capacity)
      return fail(CE_ERR_INVALID_ARGUMENT, "a mixed engine takes all its sims at once");
  } else if (n_sims > e->capacity) {
    return fail(CE_ERR_INVALID_ARGUMENT, "n_sims exceeds capacity");
  }
  int rc = ce_engine_enable_state(e, 1);
  if (rc != CE_OK) return rc;
  CE_HIP(hipSetDevice(e->device));
  if (!e->mixed) e->sp[0].count = n_sims;
  size_t n_pos = 0;
  for (int s = 0; s < e->n_species; ++s) n_pos += e->sp[s].count * 3 * e->sp[s].n;
  const bool f32 = precision == CE_PRECISION_F32;
  const size_t el = f32 ? sizeof(float) : sizeof(double);
  const double* src[2] = {pos, vel};
  void* dst[2] = {e->state_x, e->state_v};
  for (int b = 0; b < 2; ++b) {
    std::vector<char> raw(n_pos * el);
    for (size_t i = 0; i < n_pos; ++i) {
      if (f32)
        reinterpret_cast<float*>(raw.data())[i] = static_cast<float>(src[b][i]);
      else
        reinterpret_cast<double*>(raw.data())[i] = src[b][i];
    }
    rc = copy_per_atom(e, raw.data(), dst[b], el, 3, hipMemcpyHostToDevice);
    if (rc != CE_OK) return rc;
    CE_HIP(hipStreamSynchronize(e->stream));
  }
  std::vector<double> zero(e->loaded(), 0.0);
  rc = copy_per_sim(e, const_cast<double*>(h ? h : zero.data()), e->state_h, sizeof(double),
                    hipMemcpyHostToDevice);
  if (rc == CE_OK)
    rc = copy_per_sim(e, const_cast<double*>(e0 ? e0 : zero.data()), e->state_e0, sizeof(double),
                      hipMemcpyHostToDevice);
  if (rc == CE_OK) rc = copy_per_sim(e, zero.data(), e->tfin, sizeof(double), hipMemcpyHostToDevice);
  std::vector<uint32_t> zero32(e->loaded(), 0u);
  if (rc == CE_OK) rc = copy_per_sim(e, zero32.data(), e->steps, sizeof(uint32_t), hipMemcpyHostToDevice);
  if (rc == CE_OK) rc = copy_per_sim(e, zero32.data(), e->attempts, sizeof(uint32_t), hipMemcpyHostToDevice);
  const hipError_t se = hipStreamSynchronize(e->stream);
  if (rc != CE_OK) return rc;
  if (se != hipSuccess) return fail(CE_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
  e->state_prec = precision;
  e->sampled = false;
  e->have_results = false;
  return CE_OK;
}

int ce_engine_get_state(ce_engine* e, double* pos, double* vel, double* h, double* e0,
                        size_t n_sims) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (e->state_prec < 0) return fail(CE_ERR_RUNTIME, "no saved state: enable it and run first");
  if (n_sims != e->loaded()) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims must equal the loaded sims");
  CE_HIP(hipSetDevice(e->device));
  size_t n_pos = 0;
  for (int s = 0; s < e->n_species; ++s) n_pos += e->sp[s].count * 3 * e->sp[s].n;
  const bool f32 = e->state_prec == CE_PRECISION_F32;
  const size_t el = f32 ? sizeof(float) : sizeof(double);
  double* dst[2] = {pos, vel};
  void* src[2] = {e->state_x, e->state_v};
  for (int b = 0; b < 2; ++b) {
    if (dst[b] == nullptr) continue;
    std::vector<char> raw(n_pos * el);
    int rc = copy_per_atom(e, raw.data(), src[b], el, 3, hipMemcpyDeviceToHost);
    if (rc != CE_OK) return rc;
    CE_HIP(hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < n_pos; ++i)
      dst[b][i] = f32 ? static_cast<double>(reinterpret_cast<const float*>(raw.data())[i])
                      : reinterpret_cast<const double*>(raw.data())[i];
  }
  int rc = CE_OK;
  if (h) rc = copy_per_sim(e, h, e->state_h, sizeof(double), hipMemcpyDeviceToHost);
  if (rc == CE_OK && e0) rc = copy_per_sim(e, e0, e->state_e0, sizeof(double), hipMemcpyDeviceToHost);
  if (rc != CE_OK) return rc;
  CE_HIP(hipStreamSynchronize(e->stream));
  return CE_OK;
}

}  // extern "C"

namespace ce {

// The run configuration, tolerances and launch tunables of a run, shared by the
// engine's runs and ce_run_device_noalloc.
void fill_run_args(ce::RunArgs* a, const ce_run_config* cfg, double rtol, double atol,
                   double coulomb_k, const ce_launch_config& lc) {
  a->stop_rule = cfg->stop_rule;
  a->rtol = rtol;
  a->atol = atol;
  a->max_dt = cfg->max_dt;
  a->t_max = cfg->t_max;
  a->pe_stop_fraction = cfg->pe_stop_fraction;
  a->momentum_tol = cfg->momentum_tol;
  a->h0 = cfg->h0;
  a->coulomb_k = coulomb_k;
  a->max_steps = cfg->max_steps;
  a->redistribute = cfg->redistribute_energy ? 1 : 0;
  a->h0_mode = cfg->h0_mode;
  a->max_rej = cfg->max_rejections;
  a->refill_threshold = std::max(1, std::min(64, lc.refill_threshold));
  a->chunk = std::max(64, lc.chunk);
}

// Launch request of a run over the loaded sims (or, single-species engines, the
// segment-local sims of the device list `list`).
int prepare_run(ce_engine* e, const ce_run_config* cfg, const uint32_t* list, size_t n_list,
                ce::LaunchReq* req) {
  int rc = check_config(cfg);
  if (rc != CE_OK) return rc;
  if (cfg->resume && (e->state_x == nullptr || e->state_prec != cfg->precision))
    return fail(CE_ERR_RUNTIME, "resume needs a saved state from a run of the same precision");
  if (list != nullptr && e->n_species != 1)
    return fail(CE_ERR_INVALID_ARGUMENT, "sim lists need a single-species engine");
  *req = ce::LaunchReq{};
  ce::RunArgs& a = req->args;
  fill_run_args(&a, cfg, e->rtol, e->atol, e->coulomb_k, e->launch);
  a.geom = e->geom;
  a.geom_status = e->sampled ? e->gstat : nullptr;
  a.qsim = e->have_charges ? e->qsim : nullptr;
  a.ctr = e->ctr;
  a.mom = e->mom;
  a.mom32 = e->mom32;
  a.steps = e->steps;
  a.attempts = e->attempts;
  a.status = e->status;
  a.t_final = e->tfin;
  a.eerr = e->energy_diag ? e->eerr : nullptr;
  a.trace_h = e->trace;
  a.trace_cap = e->trace_cap;
  a.n_segments = e->n_species;
  a.queue_slot = 0;
  a.state_x = e->state_x;
  a.state_v = e->state_v;
  a.state_h = e->state_h;
  a.state_e0 = e->state_e0;
  a.resume = cfg->resume ? 1 : 0;
  a.sim_list = list;
  for (int s = 0; s < e->n_species; ++s) {
    uint64_t start, pos_off, q_off;
    e->offsets(s, &start, &pos_off, &q_off);
    a.seg[s] = ce::Segment{start, list != nullptr ? n_list : e->sp[s].count, pos_off, q_off,
                           e->sp[s].n, e->have_charges ? 1 : 0};
    req->species[s].n = e->sp[s].n;
    for (int i = 0; i < e->sp[s].n; ++i) {
      req->species[s].masses[i] = e->sp[s].masses[i];
      req->species[s].charges[i] = e->sp[s].charges[i];
    }
  }
  req->coulomb_k = e->coulomb_k;
  req->per_sim_charges = e->have_charges;
  req->block_threads = e->launch.block_threads;
  req->blocks_per_cu = e->launch.blocks_per_cu;
  req->mixed_launch = e->launch.mixed_launch;
  req->n_cus = e->n_cus;
  req->stream = e->stream;
  return CE_OK;
}

// Counter reset and integrator launch(es) on the engine stream: kernels only
// (no memset / copy nodes), so the sequence can be captured into a graph.
int launch_run(ce_engine* e, const ce::LaunchReq& req, int precision, size_t n_run,
               ce::LaunchInfo* info) {
  ce::clear_stale_error();
  hipLaunchKernelGGL(ce::zero_u64, dim3(1), dim3(64), 0, e->stream, e->ctr, ce::kCtrWords);
  CE_HIP(hipGetLastError());
  bool bad_n = false;
  if (n_run > 0) {
    const hipError_t he = (precision == CE_PRECISION_F64) ? ce::launch_dp5_f64(req, info, &bad_n)
                                                          : ce::launch_dp5_f32(req, info, &bad_n);
    if (bad_n) return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
    if (he != hipSuccess) return fail(CE_ERR_HIP, std::string("dp5 launch: ") + hipGetErrorString(he));
  }
  return CE_OK;
}

// After the k0 / k1 events bracketing a run: counters, timing and stats.
int finish_run(ce_engine* e, size_t n_run, const ce::LaunchInfo& info, int precision) {
  unsigned long long host_ctr[ce::kCtrWords] = {};
  CE_HIP(hipMemcpyAsync(host_ctr, e->ctr, sizeof(host_ctr), hipMemcpyDeviceToHost, e->stream));
  CE_HIP(hipStreamSynchronize(e->stream));
  float ms = 0.0f;
  CE_HIP(hipEventElapsedTime(&ms, e->k0, e->k1));
  e->stats.kernel_ms = ms;
  e->stats.sims = n_run;
  e->stats.attempted_steps = host_ctr[0];
  e->stats.accepted_steps = host_ctr[1];
  e->stats.refills = host_ctr[2];
  e->stats.lane_slots = host_ctr[3];
  e->stats.grid_blocks = info.grid;
  e->stats.block_threads = info.block;
  e->stats.n_cus = e->n_cus;
  e->stats.launches = info.launches;
  e->have_results = true;
  e->eerr_valid = e->energy_diag;
  if (e->state_x != nullptr) e->state_prec = precision;
  return CE_OK;
}

int engine_run(ce_engine* e, const ce_run_config* cfg, const uint32_t* list, size_t n_list) {
  CE_HIP(hipSetDevice(e->device));
  ce::LaunchReq req;
  int rc = prepare_run(e, cfg, list, n_list, &req);
  if (rc != CE_OK) return rc;
  const size_t n_run = list != nullptr ? n_list : e->loaded();
  ce::LaunchInfo info;
  CE_HIP(hipEventRecord(e->k0, e->stream));
  rc = launch_run(e, req, cfg->precision, n_run, &info);
  if (rc != CE_OK) return rc;
  CE_HIP(hipEventRecord(e->k1, e->stream));
  return finish_run(e, n_run, info, cfg->precision);
}

}  // namespace ce

extern "C" {

int ce_engine_capture(ce_engine* e, const ce_run_config* cfg, uint64_t seed, size_t n_sims,
                      double radius, double min_sep, uint32_t max_attempts) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (e->mixed || e->n_species != 1)
    return fail(CE_ERR_INVALID_ARGUMENT, "graph capture needs a single-species engine");
  if (!(radius > 0.0) || !(min_sep >= 0.0) || max_attempts == 0)
    return fail(CE_ERR_INVALID_ARGUMENT, "bad sampler parameters");
  if (n_sims > e->capacity) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims exceeds capacity");
  if (cfg != nullptr && cfg->resume) return fail(CE_ERR_INVALID_ARGUMENT, "a captured batch samples: no resume");
  CE_HIP(hipSetDevice(e->device));
  CE_HIP(hipStreamSynchronize(e->stream));
  if (e->gexec) CE_HIP(hipGraphExecDestroy(e->gexec));
  if (e->graph) CE_HIP(hipGraphDestroy(e->graph));
  e->gexec = nullptr;
  e->graph = nullptr;
  if (e->first_dev == nullptr) {
    const int arc = hip_alloc(&e->first_dev, 1);
    if (arc != CE_OK) return arc;
  }
  e->sp[0].count = n_sims;
  e->have_charges = false;
  e->sampled = true;
  ce::LaunchReq req;
  int rc = ce::prepare_run(e, cfg, nullptr, 0, &req);
  if (rc != CE_OK) return rc;
  CE_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  rc = sample_segment(e, 0, seed, 0, radius, min_sep, max_attempts, false, e->first_dev);
  ce::LaunchInfo info;
  if (rc == CE_OK) rc = ce::launch_run(e, req, cfg->precision, n_sims, &info);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(e->stream, &g);
  if (rc != CE_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(CE_ERR_HIP, std::string("capture: ") + hipGetErrorString(ec));
  e->graph = g;
  CE_HIP(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
  e->graph_info = info;
  e->graph_cfg = *cfg;
  e->graph_sims = n_sims;
  e->graph_eerr = e->energy_diag;
  e->have_results = false;
  return CE_OK;
}

int ce_engine_replay(ce_engine* e, uint64_t first_index, ce_results* out) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (e->gexec == nullptr) return fail(CE_ERR_RUNTIME, "nothing captured: call ce_engine_capture");
  CE_HIP(hipSetDevice(e->device));
  e->sp[0].count = e->graph_sims;
  e->have_charges = false;
  e->sampled = true;
  ce::clear_stale_error();
  hipLaunchKernelGGL(ce::set_u64, dim3(1), dim3(64), 0, e->stream, e->first_dev, first_index);
  CE_HIP(hipGetLastError());
  CE_HIP(hipEventRecord(e->k0, e->stream));
  CE_HIP(hipGraphLaunch(e->gexec, e->stream));
  CE_HIP(hipEventRecord(e->k1, e->stream));
  const int rc = ce::finish_run(e, e->graph_sims, e->graph_info, e->graph_cfg.precision);
  if (rc != CE_OK) return rc;
  e->eerr_valid = e->graph_eerr;
  e->stats.sampler_ms = 0.0;  // inside the graph, in kernel_ms
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
  if (e->n_species != 1)
    return fail(CE_ERR_INVALID_ARGUMENT, "records are written per molecule: single-species engines only");
  if (!e->have_results) return fail(CE_ERR_RUNTIME, "no results: run the engine first");
  if (first > e->loaded() || n_sims > e->loaded() - first)
    return fail(CE_ERR_INVALID_ARGUMENT, "record range exceeds the loaded sims");
  if (n_sims == 0) return CE_OK;
  CE_HIP(hipSetDevice(e->device));
  const int n = e->sp[0].n;
  const size_t rb = ce_record_bytes(n);
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
  ce::clear_stale_error();
  switch (n) {
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

int ce_engine_pack_records_async(ce_engine* e, uint64_t first, size_t n_sims, uint8_t* host_out) {
  if (e == nullptr || (host_out == nullptr && n_sims > 0))
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (e->n_species != 1)
    return fail(CE_ERR_INVALID_ARGUMENT, "records are written per molecule: single-species engines only");
  if (!e->have_results) return fail(CE_ERR_RUNTIME, "no results: run the engine first");
  if (first > e->loaded() || n_sims > e->loaded() - first)
    return fail(CE_ERR_INVALID_ARGUMENT, "record range exceeds the loaded sims");
  if (n_sims == 0) return CE_OK;
  CE_HIP(hipSetDevice(e->device));
  if (e->copy_stream == nullptr) {
    CE_HIP(hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking));
    for (int s = 0; s < 2; ++s) {
      CE_HIP(hipEventCreateWithFlags(&e->rec_packed[s], hipEventDisableTiming));
      CE_HIP(hipEventCreateWithFlags(&e->rec_copied[s], hipEventDisableTiming));
    }
  }
  const int n = e->sp[0].n;
  const size_t rb = ce_record_bytes(n);
  const int s = e->rec_slot;
  if (e->rec_async[s] == nullptr || e->rec_async_cap[s] < n_sims) {
    CE_HIP(hipStreamSynchronize(e->copy_stream));  // the buffer's last copy is done
    if (e->rec_async[s]) CE_HIP(hipFree(e->rec_async[s]));
    e->rec_async[s] = nullptr;
    e->rec_async_cap[s] = 0;
    const int rc = hip_alloc(&e->rec_async[s], n_sims * rb);
    if (rc != CE_OK) return rc;
    e->rec_async_cap[s] = n_sims;
  }
  // the pack may overwrite this buffer only once its previous copy has left
  CE_HIP(hipStreamWaitEvent(e->stream, e->rec_copied[s], 0));
  const dim3 grid(static_cast<unsigned>((n_sims + ce::kRecBlock - 1) / ce::kRecBlock));
  const dim3 block(ce::kRecBlock);
  ce::clear_stale_error();
  switch (n) {
#define CE_PACK(NN)                                                                            \
  case NN:                                                                                     \
    hipLaunchKernelGGL(ce::pack_records<NN>, grid, block, 0, e->stream, e->geom, e->mom32,     \
                       e->steps, e->status, first, static_cast<uint64_t>(n_sims),              \
                       e->rec_async[s]);                                                       \
    break;
    CE_PACK(1) CE_PACK(2) CE_PACK(3) CE_PACK(4) CE_PACK(5) CE_PACK(6) CE_PACK(7) CE_PACK(8)
    CE_PACK(9) CE_PACK(10)
#undef CE_PACK
    default: return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
  }
  CE_HIP(hipGetLastError());
  CE_HIP(hipEventRecord(e->rec_packed[s], e->stream));
  CE_HIP(hipStreamWaitEvent(e->copy_stream, e->rec_packed[s], 0));
  CE_HIP(hipMemcpyAsync(host_out, e->rec_async[s], n_sims * rb, hipMemcpyDeviceToHost,
                        e->copy_stream));
  CE_HIP(hipEventRecord(e->rec_copied[s], e->copy_stream));
  e->rec_slot = 1 - s;
  return CE_OK;
}

int ce_engine_records_wait(ce_engine* e) {
  if (e == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (e->copy_stream == nullptr) return CE_OK;
  CE_HIP(hipSetDevice(e->device));
  CE_HIP(hipStreamSynchronize(e->copy_stream));
  return CE_OK;
}

int ce_host_alloc(size_t bytes, void** out) {
  if (out == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (bytes == 0) return CE_OK;
  CE_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
  return CE_OK;
}

int ce_host_free(void* p) {
  if (p != nullptr) CE_HIP(hipHostFree(p));
  return CE_OK;
}

int ce_batched_accelerations_f32(size_t n, size_t k, const float* px, const float* py,
                                 const float* pz, const float* pair_const, const float* inv_mass,
                                 float* ax, float* ay, float* az) {
  return soa_accelerations<float>(n, k, px, py, pz, pair_const, inv_mass, ax, ay, az);
}

int ce_batched_accelerations_f64(size_t n, size_t k, const double* px, const double* py,
                                 const double* pz, const double* pair_const,
                                 const double* inv_mass, double* ax, double* ay, double* az) {
  return soa_accelerations<double>(n, k, px, py, pz, pair_const, inv_mass, ax, ay, az);
}

int ce_batched_potential_energy_f32(size_t n, size_t k, const float* px, const float* py,
                                    const float* pz, const float* pair_const, float* pe) {
  return soa_potential_energy<float>(n, k, px, py, pz, pair_const, pe);
}

int ce_batched_potential_energy_f64(size_t n, size_t k, const double* px, const double* py,
                                    const double* pz, const double* pair_const, double* pe) {
  return soa_potential_energy<double>(n, k, px, py, pz, pair_const, pe);
}

}  // extern "C"

// ---- ce_run_device_noalloc: a run with no device allocation ------------------
namespace {

// The integrator's scratch memory for ce_run_device_noalloc: the work-queue
// head and the run counters (ce::kCtrWords 64-bit words, 64 B).  Every GPU holds
// its own copy at its own address until the process ends.  In host code this
// name is a host-side stand-in, not a device address: kernels and copies take
// the address hipGetSymbolAddress returns for the current device
// (noalloc_device), never `scratch` itself.
constexpr int kScratchWords = 8;
static_assert(kScratchWords >= ce::kCtrWords, "scratch must hold the run counters");
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

struct NoallocDevice {
  unsigned long long* scratch = nullptr;  // device address of `scratch` on this GPU
  int n_cus = 0;
};
constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
NoallocDevice g_noalloc[kNoallocMaxDevices];

// Set-up of the current device, once: the device address of its `scratch` and
// its compute-unit count (the persistent kernels launch per CU).
int noalloc_device(NoallocDevice* out) {
  int dev = 0;
  CE_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoallocMaxDevices)
    return fail(CE_ERR_RUNTIME, "ce_run_device_noalloc: device id beyond 63");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  NoallocDevice& d = g_noalloc[dev];
  if (d.scratch == nullptr) {
    int cus = 0;
    CE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    void* p = nullptr;
    CE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) return fail(CE_ERR_HIP, "hipGetSymbolAddress(scratch) gave NULL");
    d.n_cus = cus;
    d.scratch = static_cast<unsigned long long*>(p);
  }
  *out = d;
  return CE_OK;
}

}  // namespace

extern "C" int ce_run_device_noalloc(const ce_molecule_desc* mol, const ce_tolerances* tol,
                                     const ce_run_config* cfg, const double* geom,
                                     size_t n_sims, const ce_device_results* out,
                                     void* stream) {
  if (mol == nullptr || tol == nullptr || out == nullptr)
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (mol->n_atoms < 1 || mol->n_atoms > CE_MAX_ATOMS)
    return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms must be in [1, CE_MAX_ATOMS]");
  if (mol->masses == nullptr || mol->charges == nullptr)
    return fail(CE_ERR_INVALID_ARGUMENT, "masses/charges NULL");
  int rc = check_species(mol->masses, mol->n_atoms);
  if (rc != CE_OK) return rc;
  if (!(tol->rtol > 0.0) || !(tol->atol >= 0.0))
    return fail(CE_ERR_INVALID_ARGUMENT, "rtol must be > 0 and atol >= 0");
  rc = check_config(cfg);
  if (rc != CE_OK) return rc;
  if (cfg->resume) return fail(CE_ERR_INVALID_ARGUMENT, "ce_run_device_noalloc keeps no state: no resume");
  if (n_sims > 0 && (geom == nullptr || out->momenta == nullptr || out->momenta_f32 == nullptr ||
                     out->steps == nullptr || out->attempts == nullptr ||
                     out->status == nullptr || out->t_final == nullptr))
    return fail(CE_ERR_INVALID_ARGUMENT, "geometry and every result pointer are required");
  if (n_sims > SIZE_MAX / (3 * CE_MAX_ATOMS * sizeof(double)))
    return fail(CE_ERR_INVALID_ARGUMENT, "n_sims too large (buffer sizes overflow)");
  if (n_sims == 0) return CE_OK;
  NoallocDevice dev;
  rc = noalloc_device(&dev);
  if (rc != CE_OK) return rc;

  ce::LaunchReq req{};
  ce::RunArgs& a = req.args;
  const ce_launch_config lc = {0, 0, 4, 64, CE_MIXED_AUTO};  // the engine's defaults
  ce::fill_run_args(&a, cfg, tol->rtol, tol->atol, mol->coulomb_k, lc);
  a.geom = geom;
  a.ctr = dev.scratch;
  a.mom = out->momenta;
  a.mom32 = out->momenta_f32;
  a.steps = out->steps;
  a.attempts = out->attempts;
  a.status = out->status;
  a.t_final = out->t_final;
  a.n_segments = 1;
  a.seg[0] = ce::Segment{0, n_sims, 0, 0, mol->n_atoms, 0};
  req.species[0].n = mol->n_atoms;
  for (int i = 0; i < mol->n_atoms; ++i) {
    req.species[0].masses[i] = mol->masses[i];
    req.species[0].charges[i] = mol->charges[i];
  }
  req.coulomb_k = mol->coulomb_k;
  req.n_cus = dev.n_cus;
  req.stream = static_cast<hipStream_t>(stream);
  // The counters start at zero on every call (a kernel, not a memset node).
  ce::clear_stale_error();
  hipLaunchKernelGGL(ce::zero_u64, dim3(1), dim3(64), 0, req.stream, dev.scratch, ce::kCtrWords);
  CE_HIP(hipGetLastError());
  ce::LaunchInfo info;
  bool bad_n = false;
  const hipError_t he = (cfg->precision == CE_PRECISION_F64) ? ce::launch_dp5_f64(req, &info, &bad_n)
                                                             : ce::launch_dp5_f32(req, &info, &bad_n);
  if (bad_n) return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
  if (he != hipSuccess) return fail(CE_ERR_HIP, std::string("dp5 launch: ") + hipGetErrorString(he));
  return CE_OK;
}

// ---- ce_hbm_copy_calibration -------------------------------------------------
namespace {

// A plain device copy: 16-byte loads and stores, grid-stride.
__global__ void __launch_bounds__(256) hbm_copy(const uint4* __restrict__ src,
                                               uint4* __restrict__ dst, size_t n16) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

}  // namespace

extern "C" int ce_hbm_copy_calibration(size_t bytes, int32_t reps, double* ms_per_copy) {
  if (bytes == 0 || bytes % 16 != 0 || reps < 1 || ms_per_copy == nullptr)
    return fail(CE_ERR_INVALID_ARGUMENT, "bytes must be a positive multiple of 16, reps >= 1");
  int dev = 0, cus = 0;
  CE_HIP(hipGetDevice(&dev));
  CE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  void *src = nullptr, *dst = nullptr;
  CE_HIP(hipMalloc(&src, bytes));
  int rc = CE_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto chk = [&](hipError_t h) {
    if (h == hipSuccess) return;
    ce::clear_stale_error();
    if (rc == CE_OK) rc = fail(CE_ERR_HIP, hipGetErrorString(h));
  };
  chk(hipMalloc(&dst, bytes));
  chk(hipMemset(src, 1, bytes));
  chk(hipEventCreate(&e0));
  chk(hipEventCreate(&e1));
  if (rc == CE_OK) {
    const size_t n16 = bytes / 16;
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((n16 + 255) / 256, 32u * cus));
    chk(hipEventRecord(e0, nullptr));
    ce::clear_stale_error();
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(hbm_copy, dim3(grid), dim3(256), 0, nullptr,
                         static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
    chk(hipGetLastError());
    chk(hipEventRecord(e1, nullptr));
    chk(hipEventSynchronize(e1));
    float ms = 0.0f;
    chk(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_copy = static_cast<double>(ms) / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(dst);
  (void)hipFree(src);
  return rc;
}
