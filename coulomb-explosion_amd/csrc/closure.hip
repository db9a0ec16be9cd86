// Learned asymptotic closure (no reference counterpart).
//
// A run integrates every sim with the plain DP5(4) up to an early cut time
// t_cut (the engine's momentum-or-t_max rule with t_max = t_cut), keeping each
// sim's state.  For the sims stopped by the cut, a feature kernel builds the
// network input and a closed-form physical prior of the asymptotic momenta; an
// ensemble of MLPs predicts the residual over the prior; a gate kernel accepts
// the ensemble mean where the members agree to within the tolerance (per-atom
// relative spread of the predicted momenta) and queues the rest, which the
// engine then continues from their saved state to the full stop rule.
//
// Prior.  With every atom moving on a straight line from the cut on
// (r_ij(tau) = r + u tau, u = v_i - v_j), the remaining Coulomb impulse of pair
// (i, j) on atom i is
//   J = k q_i q_j int_0^inf r(tau) / |r(tau)|^3 dtau
//     = k q_i q_j (r/|r| + u/|u|) / (|r| |u| + r.u),
// and the prior p_i + sum_j J_ij is rescaled so that its kinetic energy equals
// E0 (all potential energy converted), as the engine's redistribution does.
//
// Features (8n per sim): for each atom its position relative to the centre of
// mass and its momentum m v (6n), then the n charges and the n masses; each is
// shifted and scaled by the normalisation set from the training data.
// Outputs (3n): (p_final - p_prior) / out_scale per component.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "coulomb_explosion.h"
#include "engine.hpp"
#include "mlp.hpp"

namespace {

constexpr int kMaxMembers = 8;

struct ClosureConst {
  double m[CE_MAX_ATOMS];
  double q[CE_MAX_ATOMS];
  double k;
};

struct FeatureArgs {
  const float* x;       // saved state [sim][n][3]
  const float* v;
  const double* e0;
  const uint8_t* status;
  const float* qsim;    // per-sim charges [sim][n] or nullptr
  const float* shift;   // [8n] or nullptr (raw features)
  const float* inv_scale;
  uint64_t first, count;  // sims [first, first + count)
  int all;                // 1: every sim, 0: only sims stopped by the cut
  float* feat;            // [count][8n], or column-major [8n][ld]
  double* prior;          // [count][3n], or column-major [3n][ld]
  uint64_t ld;            // 0: row-major; else the column-major leading dimension
};

constexpr int kFeatBlock = 128;  // sims per block of the feature kernel

// One thread per sim.  Column-major (COL, the closure run): every thread writes
// its sim's column entries, coalesced across the block.  Row-major (the host
// feature call): the rows are staged in LDS and written back coalesced.  The
// state stays in single precision in registers (it is stored so) and each use
// widens it.  (Staging the block's saved positions and velocities through LDS
// for coalesced loads measured slower: 112 -> 178 us per 1M CH4 sims.)
template <int N, bool COL>
__global__ void __launch_bounds__(kFeatBlock) closure_features(const FeatureArgs a,
                                                               const ClosureConst cc) {
  const uint64_t r0 = blockIdx.x * uint64_t(kFeatBlock);
  const uint64_t r = r0 + threadIdx.x;
  const uint64_t g = a.first + r;
  float fs_row[COL ? 1 : 8 * N];   // unused when COL
  double ps_row[COL ? 1 : 3 * N];  // unused when COL
  const bool active = r < a.count && (a.all || a.status[g] == ce::kSimOkTmax);
  if (active) {
    float x[3 * N], v[3 * N];
    double q[N];
    double cm[3] = {0.0, 0.0, 0.0}, mt = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      q[i] = a.qsim != nullptr ? static_cast<double>(a.qsim[g * N + i]) : cc.q[i];
      mt += cc.m[i];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        x[3 * i + c] = a.x[g * 3 * N + 3 * i + c];
        v[3 * i + c] = a.v[g * 3 * N + 3 * i + c];
        cm[c] += cc.m[i] * static_cast<double>(x[3 * i + c]);
      }
    }
    // prior: straight-line impulse of every pair, then KE rescaled to E0
    double pp[3 * N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) pp[3 * i + c] = cc.m[i] * static_cast<double>(v[3 * i + c]);
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
      for (int j = i + 1; j < N; ++j) {
        double rr[3], uu[3];
        double c2 = 0.0, a2 = 0.0, b = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          rr[c] = static_cast<double>(x[3 * i + c]) - static_cast<double>(x[3 * j + c]);
          uu[c] = static_cast<double>(v[3 * i + c]) - static_cast<double>(v[3 * j + c]);
          c2 += rr[c] * rr[c];
          a2 += uu[c] * uu[c];
          b += rr[c] * uu[c];
        }
        const double rn = sqrt(c2), un = sqrt(a2);
        const double w = rn * un + b;
        if (!(rn > 0.0) || !(un > 0.0) || !(w > 1e-300)) continue;
        // s / rn and s / un with s = k q_i q_j / w, through one division
        const double kq = cc.k * q[i] * q[j] / (w * rn * un);
        const double sr = kq * un, su = kq * rn;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double J = sr * rr[c] + su * uu[c];
          pp[3 * i + c] += J;
          pp[3 * j + c] -= J;
        }
      }
    }
    double ke = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
      ke += 0.5 * (pp[3 * i] * pp[3 * i] + pp[3 * i + 1] * pp[3 * i + 1] +
                   pp[3 * i + 2] * pp[3 * i + 2]) / cc.m[i];
    const double sc = (ke > 0.0 && a.e0[g] > 0.0) ? sqrt(a.e0[g] / ke) : 1.0;
#pragma unroll
    for (int j = 0; j < 3 * N; ++j) {
      if constexpr (COL)
        a.prior[j * a.ld + r] = pp[j] * sc;
      else
        ps_row[j] = pp[j] * sc;
    }
    // features
#pragma unroll
    for (int j = 0; j < 8 * N; ++j) {
      const int i = j < 6 * N ? j / 6 : (j < 7 * N ? j - 6 * N : j - 7 * N);
      const int c = j < 6 * N ? j % 6 : 0;
      float f;
      if (j >= 7 * N)
        f = static_cast<float>(cc.m[i]);
      else if (j >= 6 * N)
        f = static_cast<float>(q[i]);
      else if (c < 3)
        f = static_cast<float>(static_cast<double>(x[3 * i + c]) - cm[c] / mt);
      else
        f = static_cast<float>(cc.m[i] * static_cast<double>(v[3 * i + c - 3]));
      if (a.shift != nullptr) f = (f - a.shift[j]) * a.inv_scale[j];
      if constexpr (COL)
        a.feat[j * a.ld + r] = f;
      else
        fs_row[j] = f;
    }
  }
  if constexpr (!COL) {
    __shared__ float fs[kFeatBlock][8 * N];
    __shared__ double ps[kFeatBlock][3 * N];
    if (active) {
#pragma unroll
      for (int j = 0; j < 8 * N; ++j) fs[threadIdx.x][j] = fs_row[j];
#pragma unroll
      for (int j = 0; j < 3 * N; ++j) ps[threadIdx.x][j] = ps_row[j];
    }
    __syncthreads();
    // the block's rows are contiguous in feat / prior (rows not computed carry
    // whatever LDS held: nothing reads them)
    const uint64_t rows = a.count - r0 < kFeatBlock ? a.count - r0 : kFeatBlock;
    const float* fsf = &fs[0][0];
    for (uint64_t e = threadIdx.x; e < rows * 8 * N; e += kFeatBlock)
      a.feat[r0 * 8 * N + e] = fsf[e];
    const double* psf = &ps[0][0];
    for (uint64_t e = threadIdx.x; e < rows * 3 * N; e += kFeatBlock)
      a.prior[r0 * 3 * N + e] = psf[e];
  }
}

// The counters, 256 bytes apart (atomics on one line serialise with each other):
// fallback sims queued, sims stopped by the cut (candidates), pre-gate
// survivors of the current chunk and of all chunks.
constexpr int kFallback = 0, kCands = 64, kSurv = 128, kSurvTotal = 192, kCounters = 256;
// Threads per block of the gate and the pre-gate: every block makes one atomic
// add per list, so larger blocks make fewer of them.  The gate holds 4N doubles
// per sim: past six atoms it takes half the block, to keep its registers.
constexpr int kGateBlock = 1024;
template <int N>
constexpr int gate_block() { return N <= 6 ? kGateBlock : kGateBlock / 2; }

// Layout of the members' outputs and the prior: entry j of row R at j * es + R * rs
// (row-major: es = 1, rs = 3n; column-major: es = the chunk size, rs = 1), member
// k's outputs ms entries after member k - 1's.
struct GateArgs {
  const double* prior;      // [chunk rows]
  const float* y;           // members [0, split) at chunk rows
  const float* y2;          // members [split, members) at compact rows
  const float* out_scale;   // [3n]
  int members;
  uint64_t first, count;
  double tol2;              // gate on (spread / |p|)^2 per atom
  uint8_t* status;          // global sim index
  double* mom;              // [sim][n][3]
  float* mom32;
  uint32_t* list;           // fallback sims
  uint32_t* counts;         // kFallback, kCands
  // Compacted rows (the survivors of closure_pregate): row r of the gate is chunk
  // row rows[r] (prior, status, members [0, split)); nullptr = the chunk itself.
  const uint32_t* rows;
  const uint32_t* count_dev;  // or nullptr: the rows are min(count, *count_dev), read here
  uint64_t es, rs, ms;
  int32_t split;
  int32_t count_cands;      // add the candidates to kCands (not when the pre-gate has)
};

// Block-aggregated append (B threads): one atomic per block on each counter (the
// fallback list and the candidate count), each flagged lane gets its slot.
// Atomics on one address serialise at the memory side: per-wave ones would bound
// the kernel when every sim falls back.
template <int B>
__device__ __forceinline__ void block_append(bool cand, bool flag, uint32_t value, uint32_t* counts,
                                             uint32_t* list) {
  constexpr int W = B / 64;
  __shared__ uint32_t wave_n[W][2];
  __shared__ uint32_t base_s;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint64_t fmask = __ballot(flag), cmask = __ballot(cand);
  if (lane == 0) {
    wave_n[wave][0] = static_cast<uint32_t>(__popcll(fmask));
    wave_n[wave][1] = static_cast<uint32_t>(__popcll(cmask));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t nf = 0, nc = 0;
    for (int w = 0; w < W; ++w) {
      const uint32_t f = wave_n[w][0];
      wave_n[w][0] = nf;  // exclusive prefix: the wave's offset in the block's slots
      nf += f;
      nc += wave_n[w][1];
    }
    base_s = nf ? atomicAdd(counts + kFallback, nf) : 0u;
    if (nc) atomicAdd(counts + kCands, nc);
  }
  __syncthreads();
  if (flag) list[base_s + wave_n[wave][0] + __popcll(fmask & ((1ull << lane) - 1))] = value;
}

template <int N>
__global__ void __launch_bounds__(gate_block<N>()) closure_gate(const GateArgs a) {
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  uint64_t count = a.count;
  if (a.count_dev != nullptr) count = min(count, static_cast<uint64_t>(*a.count_dev));
  const uint64_t rr = a.rows != nullptr && r < count ? a.rows[r] : r;  // chunk row
  const uint64_t g = a.first + rr;
  const bool cand = r < count && a.status[g] == ce::kSimOkTmax;
  bool ok = false;
  if (cand) {
    // member-major passes: the 3N loads of a member are independent of each other
    const float* yr = a.y + rr * a.rs;
    const float* yr2 = a.y2 + r * a.rs - a.split * a.ms;  // member k >= split at yr2 + k ms
    const double inv_m = a.members > 0 ? 1.0 / a.members : 0.0;
    const double inv_m1 = a.members > 1 ? 1.0 / (a.members - 1) : 1.0;
    double pm[3 * N];
#pragma unroll
    for (int j = 0; j < 3 * N; ++j) pm[j] = 0.0;
    for (int k = 0; k < a.members; ++k) {
      const float* yk = (k < a.split ? yr : yr2) + k * a.ms;
#pragma unroll
      for (int j = 0; j < 3 * N; ++j) pm[j] += yk[j * a.es];
    }
#pragma unroll
    for (int j = 0; j < 3 * N; ++j) pm[j] *= inv_m;
    double var[N];
#pragma unroll
    for (int i = 0; i < N; ++i) var[i] = 0.0;
    for (int k = 0; k < a.members; ++k) {
      const float* yk = (k < a.split ? yr : yr2) + k * a.ms;
#pragma unroll
      for (int j = 0; j < 3 * N; ++j) {
        const double d = (yk[j * a.es] - pm[j]) * a.out_scale[j];
        var[j / 3] += d * d;
      }
    }
    ok = a.members >= 2;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double p2 = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int j = 3 * i + c;
        pm[j] = a.prior[j * a.es + rr * a.rs] + pm[j] * a.out_scale[j];  // predicted momentum
        p2 += pm[j] * pm[j];
      }
      ok = ok && (var[i] * inv_m1 <= a.tol2 * p2);  // false for NaN
    }
    if (ok) {
#pragma unroll
      for (int j = 0; j < 3 * N; ++j) {
        a.mom[g * 3 * N + j] = pm[j];
        a.mom32[g * 3 * N + j] = static_cast<float>(pm[j]);
      }
      a.status[g] = ce::kSimClosure;
    }
  }
  block_append<gate_block<N>()>(a.count_cands ? cand : false, cand && !ok, static_cast<uint32_t>(g), a.counts,
               a.list);
}

// Early rejection from the first two members.  The gate accepts a sim when for
// every atom D / (M - 1) <= tol^2 |P + zbar|^2, where z_k are the members' scaled
// outputs, zbar their mean, D = sum_k |z_k - zbar|^2 and P the prior.  With the
// first two members' mean z12, D12 = |z_1 - z_2|^2 / 2 and u = |zbar - z12|, the
// parallel-axis identity and the least spread of the other M - 2 members for a
// given mean (all equal) give D >= D12 + k u^2, k = 2M / (M - 2); and
// |P + zbar| <= A + u with A = |P + z12|.  So with c = (M - 1) tol^2 < k,
// acceptance needs D12 <= max_u [c (A + u)^2 - k u^2] = c k A^2 / (k - c), a bound
// the other members attain (all at one point along P + z12).  A sim past it on
// some atom is rejected here, without the other members (at the 1e-5 gate:
// nearly every sim); the others go to a compacted list for the remaining members
// and the full gate.  A relative margin of 1e-9 keeps rounding (FP64 here) on
// the side of the full gate.
struct PregateArgs {
  const double* prior;      // entry j of row R at j es + R rs (as GateArgs)
  const float* y;           // members 0 and 1, ms apart
  uint64_t es, rs, ms;
  const float* out_scale;   // [3n]
  uint64_t first, count;
  double c;                 // (members - 1) tol^2, < 1
  double k;                 // 2 members / (members - 2)
  uint8_t* status;
  uint32_t* list;           // fallback sims
  uint32_t* rows;           // survivors: chunk rows
  uint32_t* counts;         // kFallback, kCands, kSurv
};

template <int N>
__global__ void __launch_bounds__(kGateBlock) closure_pregate(const PregateArgs a) {
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const uint64_t g = a.first + r;
  const bool cand = r < a.count && a.status[g] == ce::kSimOkTmax;
  bool reject = false;
  if (cand) {
    const float* y0 = a.y + r * a.rs;
    const float* y1 = y0 + a.ms;
    const double bound = a.c * a.k / (a.k - a.c) * (1.0 + 1e-9);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double d12 = 0.0, a2 = 0.0;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const int j = 3 * i + cc;
        const double z0 = static_cast<double>(y0[j * a.es]) * a.out_scale[j];
        const double z1 = static_cast<double>(y1[j * a.es]) * a.out_scale[j];
        const double d = z0 - z1;
        const double m = a.prior[j * a.es + r * a.rs] + 0.5 * (z0 + z1);
        d12 += 0.5 * d * d;
        a2 += m * m;
      }
      reject = reject || d12 > bound * a2;  // false for NaN: left to the full gate
    }
  }
  // two lists: the rejected sims (fallback) and the surviving rows
  constexpr int W = kGateBlock / 64;
  __shared__ uint32_t wn[W][3];
  __shared__ uint32_t base_f, base_s;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const bool surv = cand && !reject, rej = cand && reject;
  const uint64_t fm = __ballot(rej), sm = __ballot(surv), cm = __ballot(cand);
  if (lane == 0) {
    wn[wave][0] = static_cast<uint32_t>(__popcll(fm));
    wn[wave][1] = static_cast<uint32_t>(__popcll(sm));
    wn[wave][2] = static_cast<uint32_t>(__popcll(cm));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t nf = 0, ns = 0, nc = 0;
    for (int w = 0; w < W; ++w) {
      const uint32_t f = wn[w][0], sv = wn[w][1];
      wn[w][0] = nf;
      wn[w][1] = ns;
      nf += f;
      ns += sv;
      nc += wn[w][2];
    }
    base_f = nf ? atomicAdd(a.counts + kFallback, nf) : 0u;
    base_s = ns ? atomicAdd(a.counts + kSurv, ns) : 0u;
    if (nc) atomicAdd(a.counts + kCands, nc);
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1;
  if (rej) a.list[base_f + wn[wave][0] + __popcll(fm & below)] = static_cast<uint32_t>(g);
  if (surv) a.rows[base_s + wn[wave][1] + __popcll(sm & below)] = static_cast<uint32_t>(r);
}

__global__ void zero_u32(uint32_t* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

__global__ void zero_counts(uint32_t* p) {
  if (threadIdx.x < kCounters && blockIdx.x == 0) p[threadIdx.x] = 0u;
}

// Total of the chunks' survivors, read once after the last chunk.
__global__ void add_u32(uint32_t* total, const uint32_t* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *total += *p;
}

}  // namespace

struct ce_closure {
  int n = 0;
  ClosureConst cc{};
  int device = 0;
  std::vector<ce_mlp*> members;
  float* shift = nullptr;      // [8n]
  float* inv_scale = nullptr;  // [8n]
  float* out_scale = nullptr;  // [3n]
  bool normalized = false;
  int64_t chunk_cap = 0;
  float* feat = nullptr;
  double* prior = nullptr;
  float* y = nullptr;          // [kMaxMembers][chunk][3n]
  uint32_t* counts = nullptr;  // [kCounters]: kFallback, kCands, kSurv, kSurvTotal
  uint32_t* rows = nullptr;    // pre-gate survivors: chunk rows [chunk]
  float* yc = nullptr;         // members 2.. on those rows [kMaxMembers - 2][chunk][3n]
  hipEvent_t ev[4] = {};
  ce_closure_stats stats{};
};

namespace {

int fail(int code, const std::string& m) { return ce::set_error(code, m); }

#define CL_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ce::clear_stale_error(); /* reported here: not left for the next check */   \
      return fail(CE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                             \
  } while (0)

template <class T>
int dalloc(T** p, int64_t n) {
  *p = nullptr;
  CL_HIP(hipMalloc(reinterpret_cast<void**>(p), (n > 0 ? n : 1) * sizeof(T)));
  return CE_OK;
}

void free_closure(ce_closure* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  void* b[] = {c->shift, c->inv_scale, c->out_scale, c->feat, c->prior, c->y, c->counts,
               c->rows, c->yc};
  for (void* p : b)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : c->ev)
    if (e) (void)hipEventDestroy(e);
  delete c;
}

int ensure_chunk(ce_closure* c, int64_t n) {
  if (n <= c->chunk_cap) return CE_OK;
  void* b[] = {c->feat, c->prior, c->y, c->rows, c->yc};
  for (void* p : b)
    if (p) CL_HIP(hipFree(p));
  c->feat = c->y = c->yc = nullptr;
  c->prior = nullptr;
  c->rows = nullptr;
  c->chunk_cap = 0;
  int rc = dalloc(&c->feat, n * 8 * c->n);
  if (rc == CE_OK) rc = dalloc(&c->prior, n * 3 * c->n);
  if (rc == CE_OK) rc = dalloc(&c->y, int64_t(kMaxMembers) * n * 3 * c->n);
  if (rc == CE_OK) rc = dalloc(&c->rows, n);
  if (rc == CE_OK) rc = dalloc(&c->yc, int64_t(kMaxMembers - 2) * n * 3 * c->n);
  if (rc == CE_OK) c->chunk_cap = n;
  return rc;
}

unsigned blocks(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }
unsigned gate_blocks(uint64_t n, int b) { return static_cast<unsigned>((n + b - 1) / b); }

int launch_features(ce_closure* c, const FeatureArgs& a, hipStream_t s) {
  ce::clear_stale_error();
  switch (c->n) {
#define CL_F(NN)                                                                         \
  case NN: {                                                                             \
    auto kern = a.ld > 0 ? closure_features<NN, true> : closure_features<NN, false>;     \
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>((a.count + kFeatBlock - 1) / kFeatBlock)), \
                       dim3(kFeatBlock), 0, s, a, c->cc);                                \
    break;                                                                               \
  }
    CL_F(1) CL_F(2) CL_F(3) CL_F(4) CL_F(5) CL_F(6) CL_F(7) CL_F(8) CL_F(9) CL_F(10)
#undef CL_F
    default: return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
  }
  CL_HIP(hipGetLastError());
  return CE_OK;
}

int launch_gate(ce_closure* c, const GateArgs& a, hipStream_t s) {
  ce::clear_stale_error();
  switch (c->n) {
#define CL_G(NN)                                                                          \
  case NN:                                                                                \
    hipLaunchKernelGGL(closure_gate<NN>, dim3(gate_blocks(a.count, gate_block<NN>())),        \
                       dim3(gate_block<NN>()), 0, s, a);                                      \
    break;
    CL_G(1) CL_G(2) CL_G(3) CL_G(4) CL_G(5) CL_G(6) CL_G(7) CL_G(8) CL_G(9) CL_G(10)
#undef CL_G
    default: return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
  }
  CL_HIP(hipGetLastError());
  return CE_OK;
}

int launch_pregate(ce_closure* c, const PregateArgs& a, hipStream_t s) {
  ce::clear_stale_error();
  switch (c->n) {
#define CL_P(NN)                                                                             \
  case NN:                                                                                   \
    hipLaunchKernelGGL(closure_pregate<NN>, dim3(gate_blocks(a.count, kGateBlock)),        \
                       dim3(kGateBlock), 0, s, a);                                            \
    break;
    CL_P(1) CL_P(2) CL_P(3) CL_P(4) CL_P(5) CL_P(6) CL_P(7) CL_P(8) CL_P(9) CL_P(10)
#undef CL_P
    default: return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms out of range");
  }
  CL_HIP(hipGetLastError());
  return CE_OK;
}

int check_engine(const ce_closure* c, const ce_engine* e) {
  if (e->n_species != 1 || e->sp[0].n != c->n)
    return fail(CE_ERR_INVALID_ARGUMENT, "the closure needs a single-species engine of its molecule size");
  return CE_OK;
}

FeatureArgs feature_args(const ce_closure* c, const ce_engine* e, uint64_t first, uint64_t count,
                         int all, bool normalize, uint64_t ld) {
  FeatureArgs a{};
  a.x = static_cast<const float*>(e->state_x);
  a.v = static_cast<const float*>(e->state_v);
  a.e0 = e->state_e0;
  a.status = e->status;
  a.qsim = e->have_charges ? e->qsim : nullptr;
  a.shift = normalize ? c->shift : nullptr;
  a.inv_scale = normalize ? c->inv_scale : nullptr;
  a.first = first;
  a.count = count;
  a.all = all;
  a.feat = c->feat;
  a.prior = c->prior;
  a.ld = ld;
  return a;
}

}  // namespace

extern "C" {

int ce_closure_create(int32_t n_atoms, const double* masses, const double* charges,
                      double coulomb_k, ce_closure** out) {
  if (out == nullptr || masses == nullptr || charges == nullptr)
    return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (n_atoms < 2 || n_atoms > CE_MAX_ATOMS)
    return fail(CE_ERR_INVALID_ARGUMENT, "n_atoms must be in [2, CE_MAX_ATOMS]");
  for (int i = 0; i < n_atoms; ++i)
    if (!(masses[i] > 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "masses must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(CE_ERR_NO_DEVICE, "no HIP device");
  ce_closure* c = new ce_closure();
  c->n = n_atoms;
  for (int i = 0; i < n_atoms; ++i) {
    c->cc.m[i] = masses[i];
    c->cc.q[i] = charges[i];
  }
  c->cc.k = coulomb_k;
  int rc = CE_OK;
  if (hipGetDevice(&c->device) != hipSuccess) rc = fail(CE_ERR_HIP, "hipGetDevice");
  if (rc == CE_OK) rc = dalloc(&c->shift, 8 * n_atoms);
  if (rc == CE_OK) rc = dalloc(&c->inv_scale, 8 * n_atoms);
  if (rc == CE_OK) rc = dalloc(&c->out_scale, 3 * n_atoms);
  if (rc == CE_OK) rc = dalloc(&c->counts, kCounters);
  for (int i = 0; i < 4 && rc == CE_OK; ++i)
    if (hipEventCreate(&c->ev[i]) != hipSuccess) rc = fail(CE_ERR_HIP, "hipEventCreate");
  if (rc != CE_OK) {
    free_closure(c);
    return rc;
  }
  *out = c;
  return CE_OK;
}

int ce_closure_destroy(ce_closure* c) {
  free_closure(c);
  return CE_OK;
}

int ce_closure_set_normalization(ce_closure* c, const float* in_shift, const float* in_scale,
                                 const float* out_scale) {
  if (!c || !in_shift || !in_scale || !out_scale) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  const int f = 8 * c->n, o = 3 * c->n;
  std::vector<float> inv(f);
  for (int j = 0; j < f; ++j) {
    if (!(in_scale[j] > 0.0f)) return fail(CE_ERR_INVALID_ARGUMENT, "input scales must be positive");
    inv[j] = 1.0f / in_scale[j];
  }
  for (int j = 0; j < o; ++j)
    if (!(out_scale[j] > 0.0f)) return fail(CE_ERR_INVALID_ARGUMENT, "output scales must be positive");
  CL_HIP(hipSetDevice(c->device));
  CL_HIP(hipMemcpy(c->shift, in_shift, f * sizeof(float), hipMemcpyHostToDevice));
  CL_HIP(hipMemcpy(c->inv_scale, inv.data(), f * sizeof(float), hipMemcpyHostToDevice));
  CL_HIP(hipMemcpy(c->out_scale, out_scale, o * sizeof(float), hipMemcpyHostToDevice));
  c->normalized = true;
  return CE_OK;
}

int ce_closure_set_members(ce_closure* c, ce_mlp* const* members, int32_t n_members) {
  if (!c || (!members && n_members > 0)) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_members < 0 || n_members > kMaxMembers)
    return fail(CE_ERR_INVALID_ARGUMENT, "0..8 ensemble members");
  for (int k = 0; k < n_members; ++k) {
    const ce_mlp* m = members[k];
    if (m == nullptr) return fail(CE_ERR_INVALID_ARGUMENT, "NULL member");
    if (ce::mlp_width(m, 0) != 8 * c->n || ce::mlp_width(m, ce::mlp_layers(m)) != 3 * c->n)
      return fail(CE_ERR_INVALID_ARGUMENT, "members must map 8n features to 3n outputs");
  }
  c->members.assign(members, members + n_members);
  return CE_OK;
}

int ce_closure_features(ce_closure* c, ce_engine* e, float* features, double* prior,
                        size_t n_sims) {
  if (!c || !e) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_engine(c, e);
  if (rc != CE_OK) return rc;
  if (e->state_prec != CE_PRECISION_F32)
    return fail(CE_ERR_RUNTIME, "no f32 saved state: enable the state and run the engine first");
  if (n_sims != e->loaded()) return fail(CE_ERR_INVALID_ARGUMENT, "n_sims must equal the loaded sims");
  CL_HIP(hipSetDevice(c->device));
  const int64_t chunk = 1 << 20;
  rc = ensure_chunk(c, std::min<int64_t>(chunk, std::max<int64_t>(1, n_sims)));
  if (rc != CE_OK) return rc;
  for (uint64_t f0 = 0; f0 < n_sims; f0 += chunk) {
    const uint64_t cnt = std::min<uint64_t>(chunk, n_sims - f0);
    rc = launch_features(c, feature_args(c, e, f0, cnt, 1, false, 0), e->stream);
    if (rc != CE_OK) return rc;
    if (features)
      CL_HIP(hipMemcpyAsync(features + f0 * 8 * c->n, c->feat, cnt * 8 * c->n * sizeof(float),
                            hipMemcpyDeviceToHost, e->stream));
    if (prior)
      CL_HIP(hipMemcpyAsync(prior + f0 * 3 * c->n, c->prior, cnt * 3 * c->n * sizeof(double),
                            hipMemcpyDeviceToHost, e->stream));
    CL_HIP(hipStreamSynchronize(e->stream));
  }
  return CE_OK;
}

// Integrates to t_cut, applies the closure where the ensemble agrees and
// continues the other sims to the stop rule of `cfg`.
int ce_closure_run(ce_closure* c, ce_engine* e, const ce_run_config* cfg, double t_cut,
                   double gate_tol, ce_results* out) {
  if (!c || !e || !cfg) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  int rc = check_engine(c, e);
  if (rc != CE_OK) return rc;
  if (cfg->stop_rule != CE_STOP_MOMENTUM_OR_TMAX || cfg->precision != CE_PRECISION_F32 || cfg->resume)
    return fail(CE_ERR_INVALID_ARGUMENT, "the closure runs the f32 momentum-or-t_max rule");
  if (!(t_cut > 0.0) || !(t_cut < cfg->t_max))
    return fail(CE_ERR_INVALID_ARGUMENT, "t_cut must be in (0, t_max)");
  if (!(gate_tol >= 0.0)) return fail(CE_ERR_INVALID_ARGUMENT, "gate_tol must be >= 0");
  if (!c->members.empty() && !c->normalized)
    return fail(CE_ERR_RUNTIME, "set the normalisation before running members");
  if (e->loaded() > UINT32_MAX)
    return fail(CE_ERR_INVALID_ARGUMENT, "the fallback list indexes at most 2^32 loaded sims");
  CL_HIP(hipSetDevice(c->device));
  rc = ce_engine_enable_state(e, 1);
  if (rc != CE_OK) return rc;
  if (e->list == nullptr) {
    CL_HIP(hipMalloc(reinterpret_cast<void**>(&e->list), std::max<size_t>(1, e->capacity) * sizeof(uint32_t)));
  }
  ce_closure_stats st{};
  const uint64_t n = e->loaded();
  st.sims = n;
  hipStream_t s = e->stream;
  CL_HIP(hipEventRecord(c->ev[0], s));
  // 1. plain integration to the cut
  ce_run_config cut = *cfg;
  cut.t_max = t_cut;
  rc = ce::engine_run(e, &cut, nullptr, 0);
  if (rc != CE_OK) return rc;
  st.cut_ms = e->stats.kernel_ms;
  st.cut_attempts = e->stats.attempted_steps;
  // 2. features, ensemble and gate, chunk by chunk
  CL_HIP(hipEventRecord(c->ev[1], s));
  ce::clear_stale_error();  // zero_counts is checked with the feature launch
  hipLaunchKernelGGL(zero_counts, dim3(1), dim3(kCounters), 0, s, c->counts);
  const int M = static_cast<int>(c->members.size());
  const int64_t chunk = 1 << 20;
  rc = ensure_chunk(c, std::min<int64_t>(chunk, std::max<int64_t>(1, n)));
  if (rc != CE_OK) return rc;
  const int o = 3 * c->n;
  // With three or more members, a gate that is not wide open and the fused
  // inference kernel, the first two members decide most sims (closure_pregate);
  // the others run on the survivors, whose count stays on the device: the member
  // passes read their rows through the survivor list, the gate reads the first
  // two members at the chunk rows and the others at the compact rows.
  // With the fused kernel the features, the prior and the members' outputs are
  // column-major (chunk rows contiguous): every kernel of the phase reads and
  // writes them coalesced.
  const double c_bound = (M - 1) * gate_tol * gate_tol;
  bool fused = true;
  for (int k = 0; k < M; ++k) fused = fused && ce::mlp_fused_ready(c->members[k]);
  const bool pregate = fused && M >= 3 && c_bound < 1.0;
  for (uint64_t f0 = 0; f0 < n && rc == CE_OK; f0 += chunk) {
    const uint64_t cnt = std::min<uint64_t>(chunk, n - f0);
    rc = launch_features(c, feature_args(c, e, f0, cnt, 0, true, fused ? cnt : 0), s);
    ce::FusedIo io{};
    if (fused) io.ldx = io.ldy = static_cast<int64_t>(cnt);
    const int first_pass = pregate ? 2 : M;
    for (int k = 0; k < first_pass && rc == CE_OK; ++k)
      rc = ce::mlp_forward_device(c->members[k], c->feat, static_cast<int64_t>(cnt),
                                  c->y + static_cast<int64_t>(k) * cnt * o, s, io);
    if (rc != CE_OK) break;
    GateArgs g{};
    g.prior = c->prior;
    g.y = c->y;
    g.out_scale = c->out_scale;
    g.members = M;
    g.first = f0;
    g.count = cnt;
    g.es = fused ? cnt : 1;
    g.rs = fused ? 1 : static_cast<uint64_t>(o);
    g.ms = cnt * o;
    g.split = M;
    g.tol2 = gate_tol * gate_tol;
    g.status = e->status;
    g.mom = e->mom;
    g.mom32 = e->mom32;
    g.list = e->list;
    g.counts = c->counts;
    g.rows = nullptr;
    g.count_cands = 1;
    st.member_rows += static_cast<uint64_t>(first_pass) * cnt;
    if (pregate) {
      hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, s, c->counts + kSurv);
      PregateArgs pa{};
      pa.prior = c->prior;
      pa.y = c->y;
      pa.es = g.es;
      pa.rs = g.rs;
      pa.ms = g.ms;
      pa.out_scale = c->out_scale;
      pa.first = f0;
      pa.count = cnt;
      pa.c = c_bound;
      pa.k = 2.0 * M / (M - 2);
      pa.status = e->status;
      pa.list = e->list;
      pa.rows = c->rows;
      pa.counts = c->counts;
      rc = launch_pregate(c, pa, s);
      if (rc != CE_OK) break;
      hipLaunchKernelGGL(add_u32, dim3(1), dim3(64), 0, s, c->counts + kSurvTotal,
                         c->counts + kSurv);
      io.rows_dev = c->counts + kSurv;
      io.idx = c->rows;
      for (int k = 2; k < M && rc == CE_OK; ++k)
        rc = ce::mlp_forward_device(c->members[k], c->feat, static_cast<int64_t>(cnt),
                                    c->yc + static_cast<int64_t>(k - 2) * cnt * o, s, io);
      if (rc != CE_OK) break;
      g.y2 = c->yc;
      g.split = 2;
      g.count_dev = c->counts + kSurv;
      g.rows = c->rows;
      g.count_cands = 0;  // counted by the pre-gate
    }
    rc = launch_gate(c, g, s);
  }
  if (rc != CE_OK) return rc;
  std::vector<uint32_t> counts(kCounters, 0u);
  CL_HIP(hipMemcpyAsync(counts.data(), c->counts, kCounters * sizeof(uint32_t),
                        hipMemcpyDeviceToHost, s));
  CL_HIP(hipEventRecord(c->ev[2], s));
  CL_HIP(hipStreamSynchronize(s));
  const uint32_t n_fb = counts[kFallback];
  st.fallback = n_fb;
  st.candidates = counts[kCands];
  if (pregate) st.member_rows += static_cast<uint64_t>(M - 2) * counts[kSurvTotal];
  // 3. fallback: continue the queued sims from their saved state
  if (n_fb > 0) {
    ce_run_config full = *cfg;
    full.resume = 1;
    rc = ce::engine_run(e, &full, e->list, n_fb);
    if (rc != CE_OK) return rc;
    st.fallback_ms = e->stats.kernel_ms;
    st.fallback_attempts = e->stats.attempted_steps;
  }
  CL_HIP(hipEventRecord(c->ev[3], s));
  CL_HIP(hipStreamSynchronize(s));
  float ms = 0.0f;
  CL_HIP(hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
  st.infer_ms = ms;
  CL_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[3]));
  st.total_ms = ms;
  e->stats.sims = n;
  c->stats = st;
  return out != nullptr ? ce_engine_download(e, out) : CE_OK;
}

int ce_closure_get_stats(const ce_closure* c, ce_closure_stats* out) {
  if (!c || !out) return fail(CE_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = c->stats;
  return CE_OK;
}

}  // extern "C"
