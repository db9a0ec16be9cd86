// Pair-lane form of the persistent DP5(4) integrator for the larger molecules
// (f32, N = 8 with per-sim or uniform pair constants).  Included by
// kernels.hpp after run_segment.
//
// One simulation occupies two neighbouring lanes (2k, 2k+1): lane half h =
// lane & 1 holds atoms [h·NA, (h+1)·NA) (NA = N/2) — their positions,
// velocities and the seven stage forces — so a lane carries half the state of
// the one-lane form and the kernel fits two waves per SIMD where the one-lane
// form, at one wave per SIMD, issues VALU only every ~4.5 cycles.  The pair
// forces are split between the two lanes: each lane evaluates its own atoms'
// pairs, the cross pairs at atom-index offsets 1..(NA-1)/2 (the partner lane
// covers the opposite offsets) and, for even NA, the offset-NA/2 pairs of its
// first NA/2 atoms (the partner lane's first NA/2 are the rest), with the
// reaction on the partner's atoms sent back through one DPP exchange, and the
// offset-0 cross pairs for its own atoms only (both lanes evaluate those).  Every other quantity of the step
// is per component (the lane's own 3·NA) except the error norm, the stop test
// and the energies, which combine the two halves as a + partner(a) —
// commutative, so both lanes hold the same bits and take the same decisions.
// The arithmetic per component is the one-lane form's; the force sums run in a
// different order, so results agree with it to rounding (the f32 bar), not bit
// for bit.
#pragma once

#include "kernels.hpp"

namespace ce {

// The partner lane's value (lanes 2k <-> 2k+1, DPP quad_perm [1,0,3,2]).  Both
// lanes of a pair are always active together.  bound_ctrl is set: the swap never
// reads outside its row, so it changes no value, and with it the compiler folds
// the move into the instruction that uses it (v_sub_f32_dpp, v_add_f32_dpp): a
// partner operand costs nothing and holds no register.
__device__ __forceinline__ float xchg(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ int xchg(int v) { return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true); }
// xchg(v) as the operand of one subtraction a - partner(v): `a` is passed as the
// move's (ignored) old value, so the uses of one partner value are distinct
// calls and each folds into its own v_sub_f32_dpp instead of sharing one held
// register.
__device__ __forceinline__ float xchg_for(float v, float a) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, a),
                                                               __builtin_bit_cast(int, v), 0xB1,
                                                               0xF, 0xF, true));
}
__device__ __forceinline__ double xchg(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = static_cast<uint32_t>(xchg(static_cast<int>(static_cast<uint32_t>(b))));
  const uint32_t hi = static_cast<uint32_t>(xchg(static_cast<int>(static_cast<uint32_t>(b >> 32))));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
// Sum over the two halves, the same bits on both lanes.
template <class V>
__device__ __forceinline__ V pair_sum(V a) { return a + xchg(a); }

template <int N, int CM>
constexpr bool use_pair_lanes() {
  return (N == 8 || N == 10) && (CM == kPerSim || CM == kUniform);
}

// Charges of a lane's pair terms.  Per-sim charges: the force sums carry
// G_i = sum_j q_j r_ij / |r_ij|^3 and the per-atom factor k q_i / m_i (the lane's
// im, set when a sim starts) turns them into accelerations, as the uniform form's
// c / m does, so a pair term needs q_j / |r|^3 and q_i / |r|^3 only and no pair
// constant is formed or held.  Uniform pair constant c: the factors are 1 and c
// is in im.
template <int NA, int CM>
struct PairLaneK {
  float qo[CM == kPerSim ? NA : 1];  // q_i of the own atoms (the partner's: from its lane)
  float pc0 = 1.0f;                  // kUniform: the common pair constant
  double pcd0 = 1.0;                 // kUniform: the same in FP64
  double kd = 1.0;                   // kPerSim: the Coulomb constant
  // Charges of the sim starting at qs (its N = 2·NA per-atom charges).
  __device__ __forceinline__ void load_charges(const float* qs, int h) {
    if constexpr (CM == kPerSim) {
#pragma unroll
      for (int i = 0; i < NA; ++i) qo[i] = qs[h * NA + i];
    }
  }
  template <int J>
  __device__ __forceinline__ float q_own() const {
    if constexpr (CM == kPerSim) return qo[J];
    else return 1.0f;
  }
  template <int J>
  __device__ __forceinline__ float q_partner() const {
    if constexpr (CM == kPerSim) return xchg(qo[J]);
    else return 1.0f;
  }
  // sum q_i q_j / r -> the potential energy
  __device__ __forceinline__ float pe(float u) const {
    if constexpr (CM == kUniform) return u * pc0;
    else if constexpr (CM == kPerSim) return u * static_cast<float>(kd);
    else return u;
  }
  // FP64 pair constants (energy diagnostic): float(k q_i) q_j exactly for per-sim
  // charges, the molecule's FP64 constant otherwise.
  template <int I, int J>
  __device__ __forceinline__ double own_d() const {
    if constexpr (CM == kPerSim)
      return static_cast<double>(static_cast<float>(kd * static_cast<double>(qo[I]))) *
             static_cast<double>(qo[J]);
    else return pcd0;
  }
  template <int I, int J>
  __device__ __forceinline__ double cross_d() const {
    if constexpr (CM == kPerSim)
      return static_cast<double>(static_cast<float>(kd * static_cast<double>(qo[I]))) *
             static_cast<double>(xchg(qo[J]));
    else return pcd0;
  }
};

// One pair term: r = xa - xb; returns 1/|r|^3 and sets inv = 1/|r|.  With
// PARTNER, xb are the partner lane's coordinates (each fetched by the
// subtraction itself).
template <bool PARTNER = false>
__device__ __forceinline__ float pair_term(const float* xa, const float* xb, float (&r)[3],
                                           float& inv) {
#pragma unroll
  for (int d = 0; d < 3; ++d) r[d] = xa[d] - (PARTNER ? xchg_for(xb[d], xa[d]) : xb[d]);
  const float d2 = fmaT(r[0], r[0], fmaT(r[1], r[1], r[2] * r[2]));
  inv = rsqrtT(d2);
  return (inv * inv) * inv;
}

// Force sums G of the lane's own atoms (over all N atoms; the per-atom factor
// im applies later) and, with WITH_PE, the sim's potential energy (both halves).
template <int NA, int CM, bool WITH_PE>
__device__ __forceinline__ void pair_force(const float (&xo)[3 * NA], float (&a)[3 * NA], float& pe,
                                           const PairLaneK<NA, CM>& pk) {
  constexpr int NA3 = 3 * NA;
  float ap[NA3];  // reactions on the partner's atoms
#pragma unroll
  for (int q = 0; q < NA3; ++q) {
    a[q] = 0.0f;
    ap[q] = 0.0f;
  }
  float u = 0.0f;
  // own pairs (both atoms here)
  static_for<0, NA>([&](auto I) {
    constexpr int i = decltype(I)::value;
    static_for<i + 1, NA>([&](auto J) {
      constexpr int j = decltype(J)::value;
      const float qi = pk.template q_own<i>(), qj = pk.template q_own<j>();
      float r[3], inv;
      const float inv3 = pair_term(&xo[3 * i], &xo[3 * j], r, inv);
      const float si = qj * inv3, sj = qi * inv3;
      if constexpr (WITH_PE) u = fmaT(qi * qj, inv, u);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        a[3 * i + d] = fmaT(r[d], si, a[3 * i + d]);
        a[3 * j + d] = fmaT(-r[d], sj, a[3 * j + d]);
      }
    });
  });
  // cross pairs with the partner's atom j: own atom i at offsets 1..(NA-1)/2 both
  // sides (the reaction to the partner; the partner lane covers the opposite
  // offsets), and for even NA at offset NA/2 for own atoms i < NA/2 (the partner
  // lane's i < NA/2 are the other half of those pairs)
  auto cross = [&](auto Ic, auto Jc) {
    constexpr int i = decltype(Ic)::value, j = decltype(Jc)::value;
    const float qi = pk.template q_own<i>(), qj = pk.template q_partner<j>();
    float r[3], inv;
    const float inv3 = pair_term<true>(&xo[3 * i], &xo[3 * j], r, inv);
    const float si = qj * inv3, sj = qi * inv3;
    if constexpr (WITH_PE) u = fmaT(qi * qj, inv, u);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      a[3 * i + d] = fmaT(r[d], si, a[3 * i + d]);
      ap[3 * j + d] = fmaT(-r[d], sj, ap[3 * j + d]);
    }
  };
  static_for<1, (NA - 1) / 2 + 1>([&](auto O) {
    constexpr int off = decltype(O)::value;
    static_for<0, NA>([&](auto I) {
      constexpr int i = decltype(I)::value;
      cross(I, std::integral_constant<int, (i + off) % NA>{});
    });
  });
  if constexpr (NA % 2 == 0) {
    static_for<0, NA / 2>([&](auto I) {
      constexpr int i = decltype(I)::value;
      cross(I, std::integral_constant<int, i + NA / 2>{});
    });
  }
  // offset 0 (atom i with the partner's atom i): both lanes evaluate these, each
  // for its own atom only, with half of the pair's energy
  static_for<0, NA>([&](auto I) {
    constexpr int i = decltype(I)::value;
    const float qi = pk.template q_own<i>(), qp = pk.template q_partner<i>();
    float r[3], inv;
    const float inv3 = pair_term<true>(&xo[3 * i], &xo[3 * i], r, inv);
    const float si = qp * inv3;
    if constexpr (WITH_PE) u = fmaT(0.5f * (qi * qp), inv, u);
#pragma unroll
    for (int d = 0; d < 3; ++d) a[3 * i + d] = fmaT(r[d], si, a[3 * i + d]);
  });
#pragma unroll
  for (int q = 0; q < NA3; ++q) a[q] += xchg(ap[q]);
  if constexpr (WITH_PE) pe = pk.pe(pair_sum(u));
}

// Hairer / scipy initial step (src/integrators.cpp:197-250) over both halves.
template <int NA, int CM>
__device__ float pair_hairer(const float (&x)[3 * NA], const float (&v)[3 * NA],
                             const float (&a0)[3 * NA], const float (&im)[NA], float rtol,
                             float atol, float max_dt, const PairLaneK<NA, CM>& pk) {
  float d0s = 0.0f, d1s = 0.0f;
#pragma unroll
  for (int q = 0; q < 3 * NA; ++q) {
    const float sx = atol + rtol * absT(x[q]);
    const float sv = atol + rtol * absT(v[q]);
    float r = x[q] / sx;
    d0s += r * r;
    r = v[q] / sv;
    d0s += r * r;
    r = v[q] / sx;
    d1s += r * r;
    r = a0[q] / sv;
    d1s += r * r;
  }
  d0s = pair_sum(d0s);
  d1s = pair_sum(d1s);
  const float den = static_cast<float>(12 * NA);
  const float d0 = sqrtT(d0s / den);
  const float d1 = sqrtT(d1s / den);
  const float h0 = (d0 < 1e-5f || d1 < 1e-5f) ? 1e-6f : 0.01f * d0 / d1;
  float xs[3 * NA], a1[3 * NA];
#pragma unroll
  for (int q = 0; q < 3 * NA; ++q) xs[q] = x[q] + v[q] * h0;
  float unused;
  pair_force<NA, CM, false>(xs, a1, unused, pk);
#pragma unroll
  for (int q = 0; q < 3 * NA; ++q) a1[q] *= im[q / 3];
  float d2s = 0.0f;
#pragma unroll
  for (int q = 0; q < 3 * NA; ++q) {
    const float sx = atol + rtol * absT(x[q]);
    const float sv = atol + rtol * absT(v[q]);
    float r = (a0[q] * h0) / sx;
    d2s += r * r;
    r = (a1[q] - a0[q]) / sv;
    d2s += r * r;
  }
  d2s = pair_sum(d2s);
  const float d2 = sqrtT(d2s / den) / h0;
  const float h1 = (d1 <= 1e-15f && d2 <= 1e-15f) ? maxT(1e-6f, h0 * 1e-3f)
                                                   : powT(0.01f / maxT(d1, d2), 0.2f);
  return minT(minT(100.0f * h0, h1), max_dt);
}

// run_segment in pair-lane form (same queue, counters, statuses and outputs).
template <int N, int CM>
__device__ __forceinline__ void run_segment_pair(const RunArgs& p, const Segment& sg,
                                                 const MolConst<float, N>& mc,
                                                 unsigned long long* queue, Totals& tot) {
  static_assert(N % 2 == 0, "pair lanes split the atoms in two equal halves");
  constexpr int NA = N / 2;
  constexpr int NA3 = 3 * NA;
  constexpr int N3 = 3 * N;
  using T = float;
  const T rtol = static_cast<T>(p.rtol), atol = static_cast<T>(p.atol);
  const T max_dt = static_cast<T>(p.max_dt);
  const T inv6n = T(1) / static_cast<T>(6 * N);
  const T tol2 = static_cast<T>(p.momentum_tol * p.momentum_tol);
  const int lane = __lane_id();
  const int hf = lane & 1;           // which half of the sim's atoms this lane holds
  const bool lead = hf == 0;         // the lane that writes the sim's per-sim outputs
  constexpr uint64_t kLeadMask = 0x5555555555555555ull;

  PairLaneK<NA, CM> pk;
  if constexpr (CM == kUniform) {
    pk.pc0 = mc.pc[0];
    pk.pcd0 = mc.pcd[0];
  }
  if constexpr (CM == kPerSim) pk.kd = p.coulomb_k;
  T im[NA];  // per-atom factor of the own atoms: c/m (kUniform); k q_i / m_i (per sim)
#pragma unroll
  for (int i = 0; i < NA; ++i) im[i] = hf ? mc.im[NA + i] : mc.im[i];
  T x[NA3], v[NA3], k[7][NA3];
  T h = T(0);
  double t = 0.0;
  T e0 = T(0);
  uint32_t steps = 0, att = 0;
  int nrej = 0;
  bool rej_once = false;
  uint64_t loc = 0;
  bool active = false;
  int status = kRunning;
  uint64_t cbase = 0, cend = 0;
  bool exhausted = false;

  for (;;) {
    // ---- finalize stopped sims (both lanes of a pair together) ----
    const uint64_t stopped = __ballot(active && status != kRunning) & kLeadMask;
    const bool fin_now =
        stopped != 0 && ((__ballot(active && status == kRunning) & kLeadMask) == 0 ||
                         2 * (__popcll(stopped) + __popcll(__ballot(!active) & kLeadMask)) >=
                             p.refill_threshold);
    if (fin_now && active && status != kRunning) {
      const ColdArgs cp = cold_args();  // the finalize's fields at their point of use
      double ke = 0.0;
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
        const double md = hf ? mc.md[NA + i] : mc.md[i];
        ke += 0.5 * md * (vx * vx + vy * vy + vz * vz);
      }
      ke = pair_sum(ke);
      const double e0d = static_cast<double>(e0);
      const double s = (cp->redistribute && ke > 0.0) ? __builtin_sqrt(e0d / ke) : 1.0;
      const uint64_t base = sg.pos_off + loc * N3 + hf * NA3;
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const double ms = (hf ? mc.md[NA + i] : mc.md[i]) * s;
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
        for (int q = 0; q < NA3; ++q) {
          sx[q] = x[q];
          sv[q] = v[q];
        }
        if (lead) {
          cp->state_h[gs] = maxT(static_cast<double>(h), cp->state_h[gs]);
          cp->state_e0[gs] = e0d;
        }
      }
      double pe_end = 0.0;
      if (cp->eerr != nullptr) {
        // energy diagnostic (KE + PE - E0) / E0 in FP64, both halves
        double xpd[NA3];
#pragma unroll
        for (int q = 0; q < NA3; ++q) xpd[q] = static_cast<double>(xchg(x[q]));
        const auto term = [&](const double* xa, const double* xb, double c, double w) {
          const double rx = xa[0] - xb[0], ry = xa[1] - xb[1], rz = xa[2] - xb[2];
          pe_end = fma(w * c, rsqrt_f64(rx * rx + ry * ry + rz * rz), pe_end);
        };
        double xod[NA3];
#pragma unroll
        for (int q = 0; q < NA3; ++q) xod[q] = static_cast<double>(x[q]);
        static_for<0, NA>([&](auto I) {
          constexpr int i = decltype(I)::value;
          static_for<i + 1, NA>([&](auto J) {
            constexpr int j = decltype(J)::value;
            term(&xod[3 * i], &xod[3 * j], pk.template own_d<i, j>(), 1.0);
          });
          static_for<0, NA>([&](auto J) {
            constexpr int j = decltype(J)::value;
            // each cross pair is seen by both lanes: half of it here
            term(&xod[3 * i], &xpd[3 * j], pk.template cross_d<i, j>(), 0.5);
          });
        });
        pe_end = pair_sum(pe_end);
      }
      if (lead) {
        tot.att += cp->resume ? att - cp->attempts[gs] : att;
        cp->steps[gs] = steps;
        cp->attempts[gs] = att;
        cp->status[gs] = static_cast<uint8_t>(status);
        cp->t_final[gs] = t;
        if (cp->eerr != nullptr) cp->eerr[gs] = (e0d != 0.0) ? (ke + pe_end - e0d) / e0d : 0.0;
        if (cp->trace_h != nullptr)
          for (uint32_t q = steps; q < cp->trace_cap; ++q) cp->trace_h[gs * cp->trace_cap + q] = 0.0;
      }
      active = false;
    }

    // ---- refill idle pairs from the work queue (wave-uniform control) ----
    const uint64_t idle = __ballot(!active) & kLeadMask;
    const int n_idle = __popcll(idle);
    if (!exhausted && 2 * n_idle >= p.refill_threshold) {
      ++tot.refills;
      // rank of this pair among the idle pairs (both lanes: the lead lane's)
      const int rank = __popcll(idle & ((1ull << (lane & ~1)) - 1));
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
        t = 0.0;
        status = kRunning;
        const uint64_t own_off = sg.pos_off + loc * N3 + hf * NA3;
        if (p.resume) {
          const T* sx = static_cast<const T*>(p.state_x) + own_off;
          const T* sv = static_cast<const T*>(p.state_v) + own_off;
#pragma unroll
          for (int q = 0; q < NA3; ++q) {
            x[q] = sx[q];
            v[q] = sv[q];
          }
        } else {
          const double* g = p.geom + own_off;
#pragma unroll
          for (int q = 0; q < NA3; ++q) {
            x[q] = static_cast<T>(g[q]);
            v[q] = T(0);
          }
        }
        if constexpr (CM == kPerSim) {
          pk.load_charges(p.qsim + sg.q_off + loc * N, hf);
#pragma unroll
          for (int i = 0; i < NA; ++i)
            im[i] = static_cast<T>(p.coulomb_k * static_cast<double>(pk.qo[i]) /
                                   (hf ? mc.md[NA + i] : mc.md[i]));
        }
        const uint64_t gs = sg.start + loc;
        if (p.resume) {
          steps = p.steps[gs];
          att = p.attempts[gs];
          t = p.t_final[gs];
          h = static_cast<T>(p.state_h[gs]);
          e0 = static_cast<T>(p.state_e0[gs]);
          T unused;
          pair_force<NA, CM, false>(x, k[0], unused, pk);
          if (steps >= p.max_steps) status = kSimNotConverged;
          if (!(h > T(0))) {  // a given state without a step size: as a fresh sim
            if (p.h0_mode == kH0Fixed) {
              h = static_cast<T>(p.h0);
            } else {
              T a0[NA3];
#pragma unroll
              for (int q = 0; q < NA3; ++q) a0[q] = k[0][q] * im[q / 3];
              h = pair_hairer<NA, CM>(x, v, a0, im, rtol, atol, max_dt, pk);
            }
          }
        } else if (p.geom_status != nullptr && p.geom_status[gs] != 0) {
          status = kSimSamplerFail;
          e0 = T(0);
        } else {
          T pe0;
          pair_force<NA, CM, true>(x, k[0], pe0, pk);
          e0 = pe0;
          if (p.stop_rule == kStopPE &&
              static_cast<double>(e0) <= p.pe_stop_fraction * static_cast<double>(e0)) {
            status = kSimOkPE;
          } else if (p.max_steps == 0) {
            status = kSimNotConverged;
          } else if (p.h0_mode == kH0Fixed) {
            h = static_cast<T>(p.h0);
          } else {
            T a0[NA3];
#pragma unroll
            for (int q = 0; q < NA3; ++q) a0[q] = k[0][q] * im[q / 3];
            h = pair_hairer<NA, CM>(x, v, a0, im, rtol, atol, max_dt, pk);
          }
        }
        if (p.state_x != nullptr && lead) p.state_h[gs] = 0.0;  // (read above on a resume)
      }
    }
    if (__ballot(active) == 0) break;
    ++tot.iters;
    tot.slots += 32;  // one seat per lane pair

    // ---- one attempted DP5(4) step on every running pair ----
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
      // the proposal of a shortened attempt to the saved-step slot (as in the
      // one-lane form; the lead lane keeps the sim's per-sim values)
      if (p.state_x != nullptr && lead && by_time && h > cap)
        p.state_h[sg.start + loc] = static_cast<double>(h);
      const T hh2 = hh * hh;
      T g[NA];
#pragma unroll
      for (int i = 0; i < NA; ++i) g[i] = hh2 * im[i];
      ++att;
      T xs[NA3], vn[NA3];
      static_for<1, 6>([&](auto Sc) {
        constexpr int s = decltype(Sc)::value;
        const T hc = hh * static_cast<T>(kDP5.c[s]);
#pragma unroll
        for (int q = 0; q < NA3; ++q) {
          const T base = fmaT(hc, v[q], x[q]);
          if constexpr (has_terms<CoefAA<s>, s>()) {
            const T acc = stage_sum<T, s>(k, q, CoefAA<s>{});
            xs[q] = fmaT(g[q / 3], acc, base);
          } else {
            xs[q] = base;
          }
        }
        T unused;
        pair_force<NA, CM, false>(xs, k[s], unused, pk);
      });
      T exs = T(0), evs = T(0);
      T evp[NA3];
      T gv[NA];
#pragma unroll
      for (int i = 0; i < NA; ++i) gv[i] = hh * im[i];
#pragma unroll
      for (int q = 0; q < NA3; ++q) {
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
      // tests after an accepted step
      T pe_new;
      pair_force<NA, CM, true>(xs, k[6], pe_new, pk);
#pragma unroll
      for (int q = 0; q < NA3; ++q) {
        const T ev = fmaT(static_cast<T>(kDP5.e[6]), k[6][q], evp[q]);
        const T scv = fmaT(rtol, max_abs(v[q], vn[q]), atol);
        const T rv = (gv[q / 3] * ev) * rcpT(scv);
        evs = fmaT(rv, rv, evs);
      }
      const T en2 = (pair_sum(exs) + pair_sum(evs)) * inv6n;
      const T grow = T(0.9) * exp2T(T(-0.1) * log2T(en2));
      if (en2 <= T(1)) {
        bool settled = true;
        if (p.stop_rule == kStopMomentumOrTmax) {
          const auto atom_settled = [&](int i) {
            const T dx = vn[3 * i] - v[3 * i], dy = vn[3 * i + 1] - v[3 * i + 1],
                    dz = vn[3 * i + 2] - v[3 * i + 2];
            const T num = fmaT(dz, dz, fmaT(dy, dy, dx * dx));
            const T den = fmaT(vn[3 * i + 2], vn[3 * i + 2],
                               fmaT(vn[3 * i + 1], vn[3 * i + 1], vn[3 * i] * vn[3 * i]));
            return num < tol2 * den;
          };
          // own atom 0 first, the rest only where some pair of the wave passes it
          int ok = atom_settled(0) ? 1 : 0;
          ok = ok & xchg(ok);
          if (__ballot(ok != 0) != 0) {
#pragma unroll
            for (int i = 1; i < NA; ++i) ok = ok & (atom_settled(i) ? 1 : 0);
            ok = ok & xchg(ok);
          }
          settled = ok != 0;
        }
#pragma unroll
        for (int q = 0; q < NA3; ++q) {
          x[q] = xs[q];
          v[q] = vn[q];
          k[0][q] = k[6][q];
        }
        T f = (en2 == T(0)) ? T(10) : minT(T(10), grow);
        if (rej_once) f = minT(f, T(1));
        // The step that lands on t_max ends the sim by the time limit; the momentum
        // test does not count on it (its change covers a shortened interval).
        const bool final_step = by_time && (hh >= cap || t + static_cast<double>(hh) >= p.t_max);
        h = minT(hh * f, final_step ? max_dt : cap);
        rej_once = false;
        nrej = 0;
        if (lead && p.trace_h != nullptr && steps < p.trace_cap)
          p.trace_h[(sg.start + loc) * p.trace_cap + steps] = static_cast<double>(hh);
        ++steps;
        if (lead) ++tot.acc;
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

}  // namespace ce
