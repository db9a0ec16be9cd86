// Device code of the engine: per-lane Coulomb force, the persistent fused
// DP5(4) integrator with device work-queue refill, and the counter-based
// uniform-sphere sampler.  gfx950, wave64.  Included by engine.hip only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "dp5_table.hpp"

namespace ce {

// ---------------------------------------------------------------- helpers ----

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float absT(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double absT(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float sqrtT(float a) { return __builtin_sqrtf(a); }
__device__ __forceinline__ double sqrtT(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float powT(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double powT(double a, double b) { return pow(a, b); }
// 1/sqrt(d2): the f32 path uses the hardware reciprocal square root
// (v_rsq_f32, 1 ulp); the f64 oracle path the correctly rounded 1/sqrt.
__device__ __forceinline__ float rsqrtT(float a) { return __builtin_amdgcn_rsqf(a); }
__device__ __forceinline__ double rsqrtT(double a) { return 1.0 / __builtin_sqrt(a); }
// 1/x inside the error norm: v_rcp_f32 (1 ulp) / exact in f64.
__device__ __forceinline__ float rcpT(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ double rcpT(double a) { return 1.0 / a; }
// 2^x and log2(x): v_exp_f32 / v_log_f32 (1 ulp) on the f32 path, the OCML
// functions on the f64 path.
__device__ __forceinline__ float exp2T(float a) { return __builtin_amdgcn_exp2f(a); }
__device__ __forceinline__ double exp2T(double a) { return exp2(a); }
__device__ __forceinline__ float log2T(float a) { return __builtin_amdgcn_logf(a); }
__device__ __forceinline__ double log2T(double a) { return log2(a); }
// std::min / std::max semantics ((b < a) ? b : a), NaN in b keeps a.
template <class T>
__device__ __forceinline__ T minT(T a, T b) { return (b < a) ? b : a; }
template <class T>
__device__ __forceinline__ T maxT(T a, T b) { return (a < b) ? b : a; }
// max(|a|, |b|) of the error scale.  The f32 path issues one v_max_f32 with abs
// source modifiers: the compiler's fmaxf would first canonicalize both inputs
// and std::max is a compare and a select.  Equal to std::max(|a|, |b|) unless an
// input is NaN, and a NaN state gives a NaN error estimate (a rejected step)
// either way.
__device__ __forceinline__ float max_abs(float a, float b) {
  float r;
  asm("v_max_f32_e64 %0, |%1|, |%2|" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double max_abs(double a, double b) { return maxT(absT(a), absT(b)); }

__device__ __forceinline__ uint64_t bcast64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// -------------------------------------------------------------- molecule ----

constexpr int pairs_of(int n) { return n * (n - 1) / 2 > 0 ? n * (n - 1) / 2 : 1; }

// Per-molecule constants, passed by value (kernel arguments -> scalar registers).
template <class T, int N>
struct MolConst {
  T pc[pairs_of(N)];       // k q_i q_j narrowed from double, pair order (0,1),(0,2)..
  T im[N];                 // 1/m_i narrowed from double
  double pcd[pairs_of(N)]; // k q_i q_j in double (energy diagnostic)
  double md[N];            // masses in double (finalize)
};

template <int N>
constexpr int pair_index(int i, int j) {  // i < j
  return i * N - i * (i + 1) / 2 + (j - i - 1);
}

// Where a launch takes its pair constants k q_i q_j from.
enum : int {
  kPairTable = 0,  // the molecule's table (kernel arguments)
  kPerSim = 1,     // each simulation's own charges
  // Every pair has the same constant c (equal charges, or N = 2): the forces are
  // summed without c, which enters through the per-atom factors (im = c / m) and
  // the potential energy (times c).  One multiply per pair less; for c = 1 the
  // results are those of the table form bit for bit.
  kUniform = 2,
};

// Pair constant k q_i q_j of pair (i, j) as the force sum uses it.
template <class T, int N, int CM>
struct PairK {
  const MolConst<T, N>& mc;
  T kq[CM == kPerSim ? N : 1];  // k q_i
  T q[CM == kPerSim ? N : 1];   // q_j
  template <int I, int J>
  __device__ __forceinline__ T get() const {
    if constexpr (CM == kPerSim)
      return kq[I] * q[J];
    else if constexpr (CM == kUniform)
      return T(1);
    else
      return mc.pc[pair_index<N>(I, J)];
  }
  // The potential energy summed with get() -> the potential energy.
  __device__ __forceinline__ T pe(T u) const {
    if constexpr (CM == kUniform)
      return u * mc.pc[0];
    else
      return u;
  }
  template <int I, int J>
  __device__ __forceinline__ double getd() const {
    if constexpr (CM == kPerSim)
      return static_cast<double>(kq[I]) * static_cast<double>(q[J]);
    else
      return mc.pcd[pair_index<N>(I, J)];
  }
};

// Accelerations a(x) of one lane (x, a: [3N] atom-major), optionally the PE.
// 1/sqrt(d) in FP64 to ~1 ulp: the hardware estimate refined by two Newton
// steps y <- y (3/2 - d/2 y^2) (each squares the relative error), for the energy
// diagnostic (a correctly rounded sqrt and division cost several times more).
__device__ __forceinline__ double rsqrt_f64(double d) {
  double y = __builtin_amdgcn_rsq(d);
  const double hd = 0.5 * d;
  y = y * fma(-hd * y, y, 1.5);
  y = y * fma(-hd * y, y, 1.5);
  return y;
}

// Pair term: r = x_i - x_j, d2 = r.r (fma), inv = 1/|r|, s = k q_i q_j inv^3,
// a_i += r s and a_j -= r s as single-rounding FMAs (Newton 3), then a_i *= 1/m_i.
template <class T, int N, int CM, bool WITH_PE, bool WITH_IM = true>
__device__ __forceinline__ void lane_force(const T (&x)[3 * N], T (&a)[3 * N], T& pe,
                                           const PairK<T, N, CM>& pk) {
#pragma unroll
  for (int q = 0; q < 3 * N; ++q) a[q] = T(0);
  T u = T(0);
  static_for<0, N>([&](auto I) {
    constexpr int i = decltype(I)::value;
    static_for<i + 1, N>([&](auto J) {
      constexpr int j = decltype(J)::value;
      const T c = pk.template get<i, j>();
      const T rx = x[3 * i] - x[3 * j];
      const T ry = x[3 * i + 1] - x[3 * j + 1];
      const T rz = x[3 * i + 2] - x[3 * j + 2];
      const T d2 = fmaT(rx, rx, fmaT(ry, ry, rz * rz));
      const T inv = rsqrtT(d2);
      const T inv3 = (inv * inv) * inv;
      const T s = c * inv3;
      if constexpr (WITH_PE) u = fmaT(c, inv, u);
      a[3 * i] = fmaT(rx, s, a[3 * i]);
      a[3 * i + 1] = fmaT(ry, s, a[3 * i + 1]);
      a[3 * i + 2] = fmaT(rz, s, a[3 * i + 2]);
      a[3 * j] = fmaT(-rx, s, a[3 * j]);
      a[3 * j + 1] = fmaT(-ry, s, a[3 * j + 1]);
      a[3 * j + 2] = fmaT(-rz, s, a[3 * j + 2]);
    });
  });
  if constexpr (WITH_IM) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      a[3 * i] *= pk.mc.im[i];
      a[3 * i + 1] *= pk.mc.im[i];
      a[3 * i + 2] *= pk.mc.im[i];
    }
  }
  if constexpr (WITH_PE) pe = pk.pe(u);
}

template <class T, int N, int CM>
__device__ __forceinline__ T lane_pe(const T (&x)[3 * N], const PairK<T, N, CM>& pk) {
  T u = T(0);
  static_for<0, N>([&](auto I) {
    constexpr int i = decltype(I)::value;
    static_for<i + 1, N>([&](auto J) {
      constexpr int j = decltype(J)::value;
      const T rx = x[3 * i] - x[3 * j];
      const T ry = x[3 * i + 1] - x[3 * j + 1];
      const T rz = x[3 * i + 2] - x[3 * j + 2];
      const T d2 = fmaT(rx, rx, fmaT(ry, ry, rz * rz));
      u = fmaT(pk.template get<i, j>(), rsqrtT(d2), u);
    });
  });
  return pk.pe(u);
}

// Sum_{l < L} coef(l) * k[l][q] with zero coefficients skipped at compile time;
// the first non-zero term is a product, the rest are FMAs in ascending l.
template <class T, int L, class CoefFn, int M>
__device__ __forceinline__ T stage_sum(const T (&k)[7][M], int q, CoefFn coef) {
  T acc = T(0);
  bool first = true;
  static_for<0, L>([&](auto Lc) {
    constexpr int l = decltype(Lc)::value;
    constexpr double c = CoefFn::template value<l>();
    if constexpr (c != 0.0) {
      if (first)
        acc = static_cast<T>(c) * k[l][q];
      else
        acc = fmaT(static_cast<T>(c), k[l][q], acc);
      first = false;
    }
  });
  (void)coef;
  return acc;
}

template <class CoefFn, int L>
constexpr bool has_terms() {
  bool any = false;
  for (int l = 0; l < L; ++l) any = any || CoefFn::value_at(l) != 0.0;
  return any;
}

template <int S>
struct CoefAA {
  template <int L>
  static constexpr double value() { return kDP5.aa[S][L]; }
  static constexpr double value_at(int l) { return kDP5.aa[S][l]; }
};
struct CoefB {
  template <int L>
  static constexpr double value() { return kDP5.b[L]; }
};
struct CoefE {
  template <int L>
  static constexpr double value() { return kDP5.e[L]; }
};
struct CoefEA {
  template <int L>
  static constexpr double value() { return kDP5.ea[L]; }
};

// Hairer / scipy initial step (src/integrators.cpp:197-250) for one lane, in
// the scalar reference's accumulation order; k0 = a(x0).
template <class T, int N, int CM>
__device__ T lane_hairer(const T (&x)[3 * N], const T (&v)[3 * N], const T (&a0)[3 * N], T rtol,
                         T atol, T max_dt, const PairK<T, N, CM>& pk) {
  T d0s = T(0), d1s = T(0);
#pragma unroll
  for (int q = 0; q < 3 * N; ++q) {
    const T sx = atol + rtol * absT(x[q]);
    const T sv = atol + rtol * absT(v[q]);
    T r = x[q] / sx;
    d0s += r * r;
    r = v[q] / sv;
    d0s += r * r;
    r = v[q] / sx;
    d1s += r * r;
    r = a0[q] / sv;
    d1s += r * r;
  }
  const T den = static_cast<T>(6 * N);
  const T d0 = sqrtT(d0s / den);
  const T d1 = sqrtT(d1s / den);
  const T h0 = (d0 < T(1e-5) || d1 < T(1e-5)) ? T(1e-6) : T(0.01) * d0 / d1;
  T xs[3 * N], a1[3 * N];
#pragma unroll
  for (int q = 0; q < 3 * N; ++q) xs[q] = x[q] + v[q] * h0;
  T unused;
  lane_force<T, N, CM, false>(xs, a1, unused, pk);
  T d2s = T(0);
#pragma unroll
  for (int q = 0; q < 3 * N; ++q) {
    const T sx = atol + rtol * absT(x[q]);
    const T sv = atol + rtol * absT(v[q]);
    T r = (a0[q] * h0) / sx;
    d2s += r * r;
    r = (a1[q] - a0[q]) / sv;
    d2s += r * r;
  }
  const T d2 = sqrtT(d2s / den) / h0;
  const T h1 = (d1 <= T(1e-15) && d2 <= T(1e-15)) ? maxT(T(1e-6), h0 * T(1e-3))
                                                   : powT(T(0.01) / maxT(d1, d2), T(0.2));
  return minT(minT(T(100) * h0, h1), max_dt);
}

// ------------------------------------------------------------- integrator ----

enum : int {
  kStopPE = 0,
  kStopMomentumOrTmax = 1,
  kH0Hairer = 0,
  kH0Fixed = 1,
};
enum : int {
  kSimOkPE = 0,
  kSimOkMomentum = 1,
  kSimOkTmax = 2,
  kSimNotConverged = 3,
  kSimRejectLimit = 4,
  kSimNonfinite = 5,
  kSimSamplerFail = 6,
  kSimClosure = 7,
  kRunning = 255,
};

// One group of simulations of the same molecule size N inside a launch.
struct Segment {
  uint64_t start;     // first global sim index
  uint64_t count;
  uint64_t pos_off;   // offset (in elements) of sim `start` in geom / mom / mom32
  uint64_t q_off;     // offset of sim `start` in qsim (per-sim charges)
  int32_t n;          // atoms
  int32_t per_sim_charges;
};
constexpr int kMaxSegments = 4;
// ctr layout: [0] attempted steps, [1] accepted steps, [2] refill events,
// [3] lane slots (simulation seats x loop iterations of every wave: 64 seats, 32
// with lane pairs), [4 + s] work-queue head
// of segment s.
constexpr int kCtrQueue = 4;
constexpr int kCtrWords = 8;

struct RunArgs {
  const double* geom;         // [sim][n][3] per segment at pos_off
  const uint8_t* geom_status; // nullptr, or non-zero = sampler failure (global sim index)
  const float* qsim;          // per-sim charges [sim][n] per segment at q_off, or nullptr
  unsigned long long* ctr;
  double* mom;                // [sim][n][3] per segment at pos_off
  float* mom32;
  uint32_t* steps;            // global sim index
  uint32_t* attempts;
  uint8_t* status;
  double* t_final;
  double* eerr;               // energy diagnostic per global sim, or nullptr (not computed)
  double* trace_h;            // [sim][trace_cap] or nullptr
  uint32_t trace_cap;
  int32_t stop_rule;
  double rtol, atol, max_dt, t_max, pe_stop_fraction, momentum_tol, h0, coulomb_k;
  uint64_t max_steps;
  int32_t redistribute;
  int32_t h0_mode;
  int32_t max_rej;
  int32_t refill_threshold;   // 1..64
  int32_t chunk;              // >= 64
  int32_t n_segments;
  int32_t queue_slot;         // single-size launches: which queue head seg[0] uses
  Segment seg[kMaxSegments];
  // Integration state at the stop (positions and velocities as T [sim][n][3] at
  // pos_off, next step size and E0 per global sim), written when state_x is set;
  // with resume = 1 every sim starts from it (and from its steps, attempts and
  // t_final) instead of from geom.
  void* state_x;
  void* state_v;
  double* state_h;
  double* state_e0;
  int32_t resume;
  int32_t pad_;
  // Single-segment launches: queue entry q integrates segment-local sim
  // sim_list[q] (seg[0].count entries), or sim q when nullptr.
  const uint32_t* sim_list;
};

// The launch's RunArgs (kernel argument 0, at offset 0 of the kernarg segment)
// read at the point of use, through an address the compiler cannot see through.
// The finalize block reads its fields this way: the output pointers and flags
// only it uses are then not held in scalar registers across the step loop, where
// they spilled to VGPR lanes and came back with v_readlane every step (CH4:
// 13.80 -> 13.59 ms).  The refill block keeps the direct reads: its scalar loads
// at the point of use measured slower (14.02 ms), it runs as often and waits on
// them before its first force.
typedef const __attribute__((address_space(4))) RunArgs* ColdArgs;
__device__ __forceinline__ ColdArgs cold_args() {
  uint64_t a = reinterpret_cast<uint64_t>(__builtin_amdgcn_kernarg_segment_ptr());
  asm volatile("" : "+s"(a));
  return reinterpret_cast<ColdArgs>(a);
}

struct Totals {
  uint32_t att = 0, acc = 0, refills = 0;
  uint32_t iters = 0;  // loop iterations of the wave (the same on every lane)
  uint64_t slots = 0;  // simulation seats x iterations (64 per iteration, 32 with lane pairs)
};

// Runs one segment to exhaustion with a persistent per-lane DP5(4): each lane
// integrates one simulation at a time with its own step size; the wave pulls
// new simulation indices from the segment's device work queue (one atomic per
// `chunk` sims) whenever at least `refill_threshold` of its lanes are idle, so
// converged lanes do not wait for stragglers.  Every loop iteration is exactly
// one attempted step for every running lane, so the wave stays convergent
// except for the short init / finalize branches.  Returns when the queue is
// exhausted and every lane of the wave has finished.
template <class T, int N, int CM, bool DEFER_K1 = false>
__device__ __forceinline__ void run_segment(const RunArgs& p, const Segment& sg,
                                            const MolConst<T, N>& mc,
                                            unsigned long long* queue, Totals& tot) {
  constexpr int N3 = 3 * N;
  const T rtol = static_cast<T>(p.rtol), atol = static_cast<T>(p.atol);
  const T max_dt = static_cast<T>(p.max_dt);
  const T inv6n = T(1) / static_cast<T>(6 * N);
  const T tol2 = static_cast<T>(p.momentum_tol * p.momentum_tol);
  const int lane = __lane_id();

  PairK<T, N, CM> pk{mc};
  T x[N3], v[N3], k[7][N3];
  T h = T(0);
  double t = 0.0;
  T e0 = T(0);  // initial PE, computed in T (so exact in T), used in FP64
  uint32_t steps = 0, att = 0;
  int nrej = 0;
  bool rej_once = false;
  // With DEFER_K1 a fresh sim's k1 comes from its first step (see the init).
  constexpr bool kDeferK1 = DEFER_K1;
  bool k1_pending = false;  // k1 of a fresh sim still to be taken from its first step
  uint64_t loc = 0;  // index of the lane's sim inside the segment
  bool active = false;
  int status = kRunning;
  uint64_t cbase = 0, cend = 0;
  bool exhausted = false;

  for (;;) {
    // ---- finalize stopped lanes: redistribution and momenta in FP64 ----
    // Batched like the refill: the (divergent) finalize runs when enough lanes
    // have stopped to refill, or when no lane is still running, instead of on
    // almost every iteration of a 64-lane wave.
    const uint64_t stopped = __ballot(active && status != kRunning);
    const bool fin_now =
        stopped != 0 && (__ballot(active && status == kRunning) == 0 ||
                         __popcll(stopped) + __popcll(__ballot(!active)) >= p.refill_threshold);
    if (fin_now && active && status != kRunning) {
      const ColdArgs cp = cold_args();
      double ke = 0.0;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
        ke += 0.5 * mc.md[i] * (vx * vx + vy * vy + vz * vz);
      }
      const double e0d = static_cast<double>(e0);
      const double s = (cp->redistribute && ke > 0.0) ? __builtin_sqrt(e0d / ke) : 1.0;
      const uint64_t base = sg.pos_off + loc * N3;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double ms = mc.md[i] * s;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double pm = ms * static_cast<double>(v[3 * i + c]);
          cp->mom[base + 3 * i + c] = pm;
          cp->mom32[base + 3 * i + c] = static_cast<float>(pm);
        }
      }
      const uint64_t gs = sg.start + loc;
      if (cp->state_x != nullptr) {
        T* sx = static_cast<T*>(cp->state_x) + base;
        T* sv = static_cast<T*>(cp->state_v) + base;
#pragma unroll
        for (int q = 0; q < N3; ++q) {
          sx[q] = x[q];
          sv[q] = v[q];
        }
        cp->state_h[gs] = maxT(static_cast<double>(h), cp->state_h[gs]);
        cp->state_e0[gs] = e0d;
      }
      tot.att += cp->resume ? att - cp->attempts[gs] : att;
      cp->steps[gs] = steps;
      cp->attempts[gs] = att;
      cp->status[gs] = static_cast<uint8_t>(status);
      cp->t_final[gs] = t;
      if (cp->eerr != nullptr) {
        // energy diagnostic (KE + PE - E0) / E0 in FP64, when asked for
        double pe_end = 0.0;
        static_for<0, N>([&](auto I) {
          constexpr int i = decltype(I)::value;
          static_for<i + 1, N>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const double rx = static_cast<double>(x[3 * i]) - static_cast<double>(x[3 * j]);
            const double ry = static_cast<double>(x[3 * i + 1]) - static_cast<double>(x[3 * j + 1]);
            const double rz = static_cast<double>(x[3 * i + 2]) - static_cast<double>(x[3 * j + 2]);
            pe_end = fma(pk.template getd<i, j>(), rsqrt_f64(rx * rx + ry * ry + rz * rz), pe_end);
          });
        });
        cp->eerr[gs] = (e0d != 0.0) ? (ke + pe_end - e0d) / e0d : 0.0;
      }
      if (cp->trace_h != nullptr)
        for (uint32_t q = steps; q < cp->trace_cap; ++q) cp->trace_h[gs * cp->trace_cap + q] = 0.0;
      active = false;
    }

    // ---- refill idle lanes from the work queue (wave-uniform control) ----
    const uint64_t idle = __ballot(!active);
    const int n_idle = __popcll(idle);
    if (!exhausted && n_idle >= p.refill_threshold) {
      ++tot.refills;
      const int rank = static_cast<int>(lanes_below(idle));
      int assigned = 0;
      uint64_t mine = ~0ull;
      while (assigned < n_idle) {
        if (cbase >= cend) {
          uint64_t b = 0;
          if (lane == 0) b = atomicAdd(queue, static_cast<unsigned long long>(p.chunk));
          b = bcast64(b);
          if (b >= sg.count) {
            exhausted = true;
            break;
          }
          cbase = b;
          cend = (b + p.chunk < sg.count) ? b + p.chunk : sg.count;
        }
        const uint64_t left = cend - cbase;
        const int take = (static_cast<uint64_t>(n_idle - assigned) < left)
                             ? (n_idle - assigned)
                             : static_cast<int>(left);
        if (!active && rank >= assigned && rank < assigned + take)
          mine = cbase + static_cast<uint64_t>(rank - assigned);
        cbase += take;
        assigned += take;
      }
      if (!active && mine != ~0ull) {
        loc = p.sim_list != nullptr ? static_cast<uint64_t>(p.sim_list[mine]) : mine;
        active = true;
        steps = 0;
        att = 0;
        nrej = 0;
        rej_once = false;
        k1_pending = false;
        t = 0.0;
        status = kRunning;
        if (p.resume) {
          const T* sx = static_cast<const T*>(p.state_x) + sg.pos_off + loc * N3;
          const T* sv = static_cast<const T*>(p.state_v) + sg.pos_off + loc * N3;
#pragma unroll
          for (int q = 0; q < N3; ++q) {
            x[q] = sx[q];
            v[q] = sv[q];
          }
        } else {
          const double* g = p.geom + sg.pos_off + loc * N3;
#pragma unroll
          for (int q = 0; q < N3; ++q) {
            x[q] = static_cast<T>(g[q]);
            v[q] = T(0);
          }
        }
        if constexpr (CM == kPerSim) {
          const float* qs = p.qsim + sg.q_off + loc * N;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            pk.q[i] = static_cast<T>(qs[i]);
            pk.kq[i] = static_cast<T>(p.coulomb_k * static_cast<double>(qs[i]));
          }
        }
        if (p.resume) {
          // Continue a stopped integration: FSAL force at the saved positions.
          const uint64_t gs = sg.start + loc;
          steps = p.steps[gs];
          att = p.attempts[gs];
          t = p.t_final[gs];
          h = static_cast<T>(p.state_h[gs]);
          e0 = static_cast<T>(p.state_e0[gs]);
          T unused;
          lane_force<T, N, CM, false, false>(x, k[0], unused, pk);
          if (steps >= p.max_steps) status = kSimNotConverged;
          if (!(h > T(0))) {
            // a given state without a step size (ce_engine_set_state) starts its
            // step size like a fresh sim
            if (p.h0_mode == kH0Fixed) {
              h = static_cast<T>(p.h0);
            } else {
              T a0[N3];
#pragma unroll
              for (int q = 0; q < N3; ++q) a0[q] = k[0][q] * mc.im[q / 3];
              h = lane_hairer<T, N, CM>(x, v, a0, rtol, atol, max_dt, pk);
            }
          }
        } else if (p.geom_status != nullptr && p.geom_status[sg.start + loc] != 0) {
          status = kSimSamplerFail;
          e0 = T(0);
        } else {
          // The stages hold forces F = m a; 1/m enters via per-atom step factors.
          if (kDeferK1 && p.h0_mode == kH0Fixed) {
            // The sim starts at rest, so the first step's second stage sits at x0
            // exactly (x0 + h c2 * 0) and its force is F(x0) to the bit: k1 is
            // taken from there, and only the potential energy is needed here.
            e0 = lane_pe<T, N, CM>(x, pk);
            k1_pending = true;
          } else {
            T pe0;
            lane_force<T, N, CM, true, false>(x, k[0], pe0, pk);
            e0 = pe0;
          }
          if (p.stop_rule == kStopPE &&
              static_cast<double>(e0) <= p.pe_stop_fraction * static_cast<double>(e0)) {
            status = kSimOkPE;
          } else if (p.max_steps == 0) {
            status = kSimNotConverged;
          } else {
            if (p.h0_mode == kH0Fixed) {
              h = static_cast<T>(p.h0);
            } else {
              T a0[N3];
#pragma unroll
              for (int q = 0; q < N3; ++q) a0[q] = k[0][q] * mc.im[q / 3];
              h = lane_hairer<T, N, CM>(x, v, a0, rtol, atol, max_dt, pk);
            }
          }
        }
        if (p.state_x != nullptr) p.state_h[sg.start + loc] = 0.0;  // (read above on a resume)
      }
    }
    if (__ballot(active) == 0) break;
    ++tot.iters;
    tot.slots += 64;

    // ---- one attempted DP5(4) step on every running lane ----
    if (active && status == kRunning) {
      T cap = max_dt;
      bool by_time = false;
      if (p.stop_rule == kStopMomentumOrTmax) {
        const double rem = p.t_max - t;
        if (rem <= p.max_dt) {
          cap = static_cast<T>(rem);
          by_time = true;
        }
      }
      const T hh = minT(h, cap);
      // With a saved state, a resume after the step that is shortened to land on
      // the time limit starts from at least the size that step was shortened
      // from (growing back from a step clamped to a sliver would take several
      // steps, each a sliver-sized momentum change): the proposal of every
      // shortened attempt goes to the sim's saved-step slot, zeroed when the sim
      // starts and combined at the stop.  Memory, not a register kept through
      // the step.
      if (p.state_x != nullptr && by_time && h > cap)
        p.state_h[sg.start + loc] = static_cast<double>(h);
      const T hh2 = hh * hh;
      T g[N];  // h^2 / m_i
#pragma unroll
      for (int i = 0; i < N; ++i) g[i] = hh2 * mc.im[i];
      ++att;
      T xs[N3], vn[N3];
      // Stages 2..6: positions x + h c_s v + (h^2/m) sum_l aa[s][l] F_l, then F(x_s).
      static_for<1, 6>([&](auto Sc) {
        constexpr int s = decltype(Sc)::value;
        const T hc = hh * static_cast<T>(kDP5.c[s]);
#pragma unroll
        for (int q = 0; q < N3; ++q) {
          const T base = fmaT(hc, v[q], x[q]);
          if constexpr (has_terms<CoefAA<s>, s>()) {
            const T acc = stage_sum<T, s>(k, q, CoefAA<s>{});
            xs[q] = fmaT(g[q / 3], acc, base);
          } else {
            xs[q] = base;
          }
        }
        T unused;
        lane_force<T, N, CM, false, false>(xs, k[s], unused, pk);
        if constexpr (kDeferK1 && s == 1) {
          // fresh sims: stage 2 was at x0, so F(x0) is the k1 their init skipped
          // (k1 is first read by stage 3)
          static_assert(!has_terms<CoefAA<1>, 1>(), "stage 2 must not read k1");
          if (__ballot(k1_pending) != 0) {
            if (k1_pending) {
#pragma unroll
              for (int q = 0; q < N3; ++q) k[0][q] = k[1][q];
            }
            k1_pending = false;
          }
        }
      });
      // Stage 7 (FSAL): the 5th-order solution; F(x_new) is the next k1.  The
      // embedded error in the RMS norm (scale = atol + rtol max(|y_old|, |y_new|),
      // the squared norm compared with 1) is summed around F(x_new): its position
      // part and the first six velocity terms need only k1..k6, which then die
      // before the last force evaluation (k7 enters the velocity error alone,
      // as the last FMA of the same ascending sum).
      static_assert(kDP5.ea[6] == 0.0 && kDP5.e[6] != 0.0 && kDP5.e[0] != 0.0);
      T exs = T(0), evs = T(0);
      T evp[N3];
      T gv[N];  // h / m_i
#pragma unroll
      for (int i = 0; i < N; ++i) gv[i] = hh * mc.im[i];
#pragma unroll
      for (int q = 0; q < N3; ++q) {
        const T ax = stage_sum<T, 6>(k, q, CoefAA<6>{});
        const T av = stage_sum<T, 6>(k, q, CoefB{});
        xs[q] = fmaT(g[q / 3], ax, fmaT(hh, v[q], x[q]));
        vn[q] = fmaT(gv[q / 3], av, v[q]);
        const T ex = stage_sum<T, 6>(k, q, CoefEA{});
        evp[q] = stage_sum<T, 6>(k, q, CoefE{});
        const T scx = fmaT(rtol, max_abs(x[q], xs[q]), atol);
        const T rx = (g[q / 3] * ex) * rcpT(scx);
        exs = fmaT(rx, rx, exs);
      }
      // F(x_new) also gives the potential energy at x_new, which the PE rule
      // tests after an accepted step: the same pair terms in the same order as
      // lane_pe, so the same bits, for one FMA per pair.
      T pe_new;
      lane_force<T, N, CM, true, false>(xs, k[6], pe_new, pk);
#pragma unroll
      for (int q = 0; q < N3; ++q) {
        const T ev = fmaT(static_cast<T>(kDP5.e[6]), k[6][q], evp[q]);
        const T scv = fmaT(rtol, max_abs(v[q], vn[q]), atol);
        const T rv = (gv[q / 3] * ev) * rcpT(scv);
        evs = fmaT(rv, rv, evs);
      }
      const T en2 = (exs + evs) * inv6n;
      // 0.9 err^(-1/5) with err = sqrt(en2): 0.9 * 2^(-0.1 log2 en2).
      const T grow = T(0.9) * exp2T(T(-0.1) * log2T(en2));
      if (en2 <= T(1)) {
        bool settled = true;
        if (p.stop_rule == kStopMomentumOrTmax) {
          // Every atom must pass; atom 0 is tested first and the others only when
          // it passes on some lane of the wave (almost never before t_max).
          const auto atom_settled = [&](int i) {
            const T dx = vn[3 * i] - v[3 * i], dy = vn[3 * i + 1] - v[3 * i + 1],
                    dz = vn[3 * i + 2] - v[3 * i + 2];
            const T num = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
            const T den = fmaT(vn[3 * i + 2], vn[3 * i + 2],
                               fmaT(vn[3 * i + 1], vn[3 * i + 1], vn[3 * i] * vn[3 * i]));
            return num < tol2 * den;
          };
          settled = atom_settled(0);
          if (__ballot(settled) != 0) {
#pragma unroll
            for (int i = 1; i < N; ++i) settled = settled && atom_settled(i);
          }
        }
#pragma unroll
        for (int q = 0; q < N3; ++q) {
          x[q] = xs[q];
          v[q] = vn[q];
          k[0][q] = k[6][q];
        }
        T f = (en2 == T(0)) ? T(10) : minT(T(10), grow);
        if (rej_once) f = minT(f, T(1));
        // The step that lands on t_max (shortened to the remaining time, or ending
        // there in FP64) ends the sim by the time limit: the momentum test measures
        // convergence only over steps the controller chose, so it does not count
        // on that one.  After it the sim is finished unless it is resumed from its
        // saved state; the step size kept for that is not clamped to the limit.
        const bool final_step = by_time && (hh >= cap || t + static_cast<double>(hh) >= p.t_max);
        h = minT(hh * f, final_step ? max_dt : cap);
        rej_once = false;
        nrej = 0;
        if (p.trace_h != nullptr && steps < p.trace_cap)
          p.trace_h[(sg.start + loc) * p.trace_cap + steps] = static_cast<double>(hh);
        ++steps;
        ++tot.acc;
        t = final_step ? p.t_max : t + static_cast<double>(hh);
        if (p.stop_rule == kStopPE) {
          if (static_cast<double>(pe_new) <= p.pe_stop_fraction * static_cast<double>(e0))
            status = kSimOkPE;
        } else if (final_step) {
          status = kSimOkTmax;
        } else if (settled) {
          status = kSimOkMomentum;
        }
        if (status == kRunning && steps >= p.max_steps) status = kSimNotConverged;
      } else {
        const T f = maxT(T(0.2), grow);
        h = hh * f;
        rej_once = true;
        if (++nrej > p.max_rej) status = (en2 != en2) ? kSimNonfinite : kSimRejectLimit;
      }
    }

  }
}

__device__ __forceinline__ void flush_totals(const RunArgs& p, Totals& tot) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    tot.att += __shfl_xor(tot.att, off);
    tot.acc += __shfl_xor(tot.acc, off);
  }
  if (__lane_id() == 0) {
    atomicAdd(p.ctr + 0, static_cast<unsigned long long>(tot.att));
    atomicAdd(p.ctr + 1, static_cast<unsigned long long>(tot.acc));
    atomicAdd(p.ctr + 2, static_cast<unsigned long long>(tot.refills));
    atomicAdd(p.ctr + 3, static_cast<unsigned long long>(tot.slots));
  }
}

}  // namespace ce

#include "kernels_pair.hpp"

namespace ce {

// Waves per SIMD the compiler must fit the f32 integrator into (its register
// budget is 512 / waves): more resident waves hide the issue latency of the
// step's dependent chains, which one or two waves per SIMD leave exposed (a
// lone wave issues a VALU instruction every ~4-5 cycles, two or more waves
// every ~2.5).  The values are the highest occupancy each size reaches without
// spilling in the step: N = 2 at six waves (80 VGPRs; five with per-sim
// charges), N = 3 at four, N = 4 and N = 5 at three (CH4: 168 VGPRs, two
// loop-invariant values spilled and reloaded only outside the step; 8% faster
// than at two waves).  N = 5 with per-sim charges and N >= 6 would spill the
// step itself at three waves (measured no faster at N = 5); the compiler's
// choice stands for them.
template <class T, int N, int CM>
constexpr int min_waves_per_simd() {
  if constexpr (!std::is_same<T, float>::value) {
    return 1;
  } else {
    constexpr int fixed[11] = {1, 1, 6, 4, 3, 3, 1, 1, 1, 1, 1};
    constexpr int per_sim[11] = {1, 1, 5, 4, 3, 1, 1, 1, 1, 1, 1};
    if constexpr (use_pair_lanes<N, CM>()) return 2;
    return CM == kPerSim ? per_sim[N] : fixed[N];
  }
}

// The integrator of one segment: the pair-lane form where it applies (f32,
// N = 8 with per-sim or uniform pair constants), else one lane per sim.
template <class T, int N, int CM, bool DEFER_K1 = false>
__device__ __forceinline__ void run_any(const RunArgs& p, const Segment& sg,
                                        const MolConst<T, N>& mc, unsigned long long* queue,
                                        Totals& tot) {
  if constexpr (std::is_same<T, float>::value && use_pair_lanes<N, CM>())
    run_segment_pair<N, CM>(p, sg, mc, queue, tot);
  else
    run_segment<T, N, CM, DEFER_K1>(p, sg, mc, queue, tot);
}

// Whether a single-size launch takes a fresh sim's k1 from its first step: for
// N = 4 and 5, where the skipped initial force outweighs the per-step check
// (CH4: 14.1 -> 13.7 ms).  Smaller molecules' forces are too cheap to pay for
// it, the larger ones lose occupancy to the flag, and in the mixed launch it
// measured slower.
template <int N>
constexpr bool defer_k1() { return N == 4 || N == 5; }

// One molecule for the whole launch (segment 0).
template <class T, int N, int CM>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 256),
                               amdgpu_waves_per_eu(min_waves_per_simd<T, N, CM>())))
dp5_persistent(const RunArgs p, const MolConst<T, N> mc) {
  Totals tot;
  run_any<T, N, CM, defer_k1<N>()>(p, p.seg[0], mc, p.ctr + kCtrQueue + p.queue_slot, tot);
  flush_totals(p, tot);
}

// Several molecule sizes in one launch: the segments are taken in order, each
// wave moving to the next segment as soon as the current one's queue is empty
// and its own lanes have finished (no grid-wide barrier between segments).  The
// sizes are compile-time (N0..N3); segment s must have size N_s.
template <class T, int N0, int N1, int N2, int N3>
struct MixedConst {
  MolConst<T, N0> m0;
  MolConst<T, N1> m1;
  MolConst<T, N2> m2;
  MolConst<T, N3> m3;
};

// f32: two waves per SIMD (the N = 8 segment's registers set the kernel's; its
// one-lane pair-table form sits at the 256-register limit of two waves).
template <class T, int N0, int N1, int N2, int N3, int CM>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 256),
                               amdgpu_waves_per_eu(std::is_same<T, float>::value ? 2 : 1)))
dp5_mixed(const RunArgs p, const MixedConst<T, N0, N1, N2, N3> mc) {
  Totals tot;
  if (p.n_segments > 0) run_any<T, N0, CM>(p, p.seg[0], mc.m0, p.ctr + kCtrQueue + 0, tot);
  if (p.n_segments > 1) run_any<T, N1, CM>(p, p.seg[1], mc.m1, p.ctr + kCtrQueue + 1, tot);
  if (p.n_segments > 2) run_any<T, N2, CM>(p, p.seg[2], mc.m2, p.ctr + kCtrQueue + 2, tot);
  if (p.n_segments > 3) run_any<T, N3, CM>(p, p.seg[3], mc.m3, p.ctr + kCtrQueue + 3, tot);
  flush_totals(p, tot);
}

// ---------------------------------------------------------------- sampler ----

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // 32 x 32 -> 64-bit products (one v_mad_u64_u32 each rather than a
    // separate low and high multiply).
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
    const uint32_t lo0 = static_cast<uint32_t>(p0), hi0 = static_cast<uint32_t>(p0 >> 32);
    const uint32_t lo1 = static_cast<uint32_t>(p1), hi1 = static_cast<uint32_t>(p1 >> 32);
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

struct SampleArgs {
  double* geom;        // [sim][n][3], this segment's first sim
  uint8_t* status;     // 0 ok / kSimSamplerFail, this segment's first sim
  float* charges;      // [sim][n] per-sim charges or nullptr
  uint64_t first_index;  // global index of the segment's first sim
  const uint64_t* first_index_dev;  // if set, added to first_index (graph replays)
  uint64_t n_sims;
  uint32_t key0, key1;
  float R, twoR, R2, min2;
  uint32_t max_attempts;
  int32_t n_atoms;
  int32_t charge_levels;   // per-sim charges uniform in {1..levels}
  int32_t max_total_charge;  // 0 = no limit
};

// One thread per simulation.  Candidate j of global sim g draws Philox4x32-10 on
// counter (j, lo32(g), hi32(g), 0): 24-bit uniforms -> fma(u, 2R, -R); kept if
// inside the ball and >= min_sep from every atom already placed
// (src/sampler.cpp:28-57 placement rule).  Per-sim charges, when asked for, use
// counter (4t + i/4, lo32(g), hi32(g), 1), word i%4, q_i = 1 + (u24 * L >> 24),
// redrawn (t+1) while their sum exceeds max_total_charge.
template <int N>
__global__ void __launch_bounds__(256) sample_sphere(const SampleArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t first =
      a.first_index + (a.first_index_dev != nullptr ? *a.first_index_dev : uint64_t(0));
  for (uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; s < a.n_sims;
       s += stride) {
    const uint64_t g = first + s;
    // One candidate per iteration, so a wave iterates until its slowest lane
    // has placed every atom (not atom by atom).  The draw sequence is the same
    // as placing atom by atom: candidate j, cube rejection (at most 4096 draws
    // per point), then the separation test against the atoms already placed.
    uint32_t cand = 0, cube = 0, tries = 0;
    int placed = 0;
    float P[3 * N];
    bool ok = true;
    while (ok && placed < N) {
      uint32_t r[4] = {cand++, static_cast<uint32_t>(g), static_cast<uint32_t>(g >> 32), 0u};
      philox4x32_10(r, a.key0, a.key1);
      float c[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float u = static_cast<float>(r[q] >> 8) * 0x1.0p-24f;
        c[q] = __builtin_fmaf(u, a.twoR, -a.R);
      }
      const float d = __builtin_fmaf(c[2], c[2], __builtin_fmaf(c[1], c[1], c[0] * c[0]));
      if (!(d <= a.R2)) {
        ok = ++cube < 4096;
        continue;
      }
      cube = 0;
      bool clears = true;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        if (j < placed) {
          const float dx = c[0] - P[3 * j], dy = c[1] - P[3 * j + 1], dz = c[2] - P[3 * j + 2];
          const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
          clears = clears && !(d2 < a.min2);
        }
      }
      if (clears) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
          if (j == placed) {
            P[3 * j] = c[0];
            P[3 * j + 1] = c[1];
            P[3 * j + 2] = c[2];
          }
        }
        ++placed;
        tries = 0;
      } else {
        ok = ++tries < a.max_attempts;
      }
    }
    if (a.charges != nullptr) {
      float Q[N];
      bool placed = false;
      for (uint32_t t = 0; t < 4096 && !placed; ++t) {
        int total = 0;
#pragma unroll
        for (int b = 0; b < (N + 3) / 4; ++b) {
          uint32_t r[4] = {4u * t + static_cast<uint32_t>(b), static_cast<uint32_t>(g),
                           static_cast<uint32_t>(g >> 32), 1u};
          philox4x32_10(r, a.key0, a.key1);
#pragma unroll
          for (int w = 0; w < 4; ++w) {
            const int i = 4 * b + w;
            if (i < N) {
              const int qi = 1 + static_cast<int>((static_cast<uint64_t>(r[w] >> 8) *
                                                   static_cast<uint32_t>(a.charge_levels)) >> 24);
              Q[i] = static_cast<float>(qi);
              total += qi;
            }
          }
        }
        placed = a.max_total_charge <= 0 || total <= a.max_total_charge;
      }
      ok = ok && placed;
      float* qo = a.charges + s * N;
#pragma unroll
      for (int i = 0; i < N; ++i) qo[i] = ok ? Q[i] : 0.0f;
    }
    double* out = a.geom + s * (3 * N);
#pragma unroll
    for (int q = 0; q < 3 * N; ++q) out[q] = ok ? static_cast<double>(P[q]) : 0.0;
    a.status[s] = ok ? 0 : kSimSamplerFail;
  }
}

// ------------------------------------------------- reference-layout kernels ----
// batched::accelerations<T> / potential_energy<T> (batched_force.hpp:31-110) on
// the reference's SoA layout p[i*k + lane], one thread per lane, with the
// reference's IEEE 1/sqrt and division so results match it bit for bit (T =
// float); T = double backs the double-precision CoulombForce of the C++ API.
template <class T, int N>
__global__ void __launch_bounds__(256) soa_accel(uint64_t k, const T* px, const T* py,
                                                 const T* pz, const T* pc, const T* im,
                                                 T* ax, T* ay, T* az) {
  const uint64_t l = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (l >= k) return;
  T x[N], y[N], z[N], fx[N], fy[N], fz[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    x[i] = px[i * k + l];
    y[i] = py[i * k + l];
    z[i] = pz[i * k + l];
    fx[i] = fy[i] = fz[i] = T(0);
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i + 1; j < N; ++j) {
      const T rx = x[i] - x[j], ry = y[i] - y[j], rz = z[i] - z[j];
      const T d2 = fmaT(rx, rx, fmaT(ry, ry, rz * rz));
      const T inv = T(1) / sqrtT(d2);
      const T inv3 = inv / d2;
      const T s = pc[i * N + j] * inv3;
      const T gx = rx * s, gy = ry * s, gz = rz * s;
      fx[i] += gx;
      fy[i] += gy;
      fz[i] += gz;
      fx[j] -= gx;
      fy[j] -= gy;
      fz[j] -= gz;
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ax[i * k + l] = fx[i] * im[i];
    ay[i * k + l] = fy[i] * im[i];
    az[i * k + l] = fz[i] * im[i];
  }
}

template <class T, int N>
__global__ void __launch_bounds__(256) soa_pe(uint64_t k, const T* px, const T* py,
                                              const T* pz, const T* pc, T* pe) {
  const uint64_t l = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (l >= k) return;
  T u = T(0);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i + 1; j < N; ++j) {
      const T rx = px[i * k + l] - px[j * k + l];
      const T ry = py[i * k + l] - py[j * k + l];
      const T rz = pz[i * k + l] - pz[j * k + l];
      const T d2 = fmaT(rx, rx, fmaT(ry, ry, rz * rz));
      u = fmaT(pc[i * N + j], T(1) / sqrtT(d2), u);
    }
  pe[l] = u;
}

// ------------------------------------------------------------ records ----
// Dataset records of examples/generate_dataset.cpp:198-217, packed little-endian
// with no padding: n*3 f32 initial positions, n*3 f32 final momenta, u32 steps,
// u8 converged -- 24n + 5 bytes (125 for CH4).  A block packs kRecBlock
// consecutive sims into LDS byte by byte, then streams the contiguous span out
// with 16-byte stores (the span starts at a multiple of 16: kRecBlock * bytes is
// a multiple of 16 because kRecBlock is).
constexpr int kRecBlock = 256;

__device__ __forceinline__ void put_bytes(uint8_t* dst, uint32_t v) {
  dst[0] = static_cast<uint8_t>(v);
  dst[1] = static_cast<uint8_t>(v >> 8);
  dst[2] = static_cast<uint8_t>(v >> 16);
  dst[3] = static_cast<uint8_t>(v >> 24);
}

template <int N>
__global__ void __launch_bounds__(kRecBlock) pack_records(const double* geom, const float* mom32,
                                                          const uint32_t* steps,
                                                          const uint8_t* status, uint64_t first,
                                                          uint64_t n_sims, uint8_t* out) {
  constexpr int kBytes = 24 * N + 5;
  __shared__ __align__(16) uint8_t stage[kRecBlock * kBytes];
  const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kRecBlock;  // relative to first
  const uint64_t s = b0 + threadIdx.x;
  if (s < n_sims) {
    const uint64_t g = first + s;
    uint8_t* r = stage + threadIdx.x * kBytes;
#pragma unroll
    for (int q = 0; q < 3 * N; ++q) put_bytes(r + 4 * q, __float_as_uint(static_cast<float>(geom[g * 3 * N + q])));
#pragma unroll
    for (int q = 0; q < 3 * N; ++q) put_bytes(r + 12 * N + 4 * q, __float_as_uint(mom32[g * 3 * N + q]));
    put_bytes(r + 24 * N, steps[g]);
    r[24 * N + 4] = status[g] <= 2 ? 1 : 0;
  }
  __syncthreads();
  const uint64_t count = (n_sims - b0 < kRecBlock) ? (n_sims - b0) : kRecBlock;
  const uint64_t bytes = count * kBytes;
  uint8_t* dst = out + b0 * kBytes;
  const uint64_t n16 = bytes / 16;
  for (uint64_t i = threadIdx.x; i < n16; i += kRecBlock)
    reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(stage)[i];
  for (uint64_t i = n16 * 16 + threadIdx.x; i < bytes; i += kRecBlock) dst[i] = stage[i];
}

static __global__ void set_u64(uint64_t* p, uint64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

static __global__ void zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace ce
