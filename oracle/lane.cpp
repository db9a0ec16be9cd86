// CPU oracle, per-lane formulation -- TEST INFRASTRUCTURE ONLY (see ce_oracle.h).
//
// The GPU engine integrates one simulation per lane with its own step size
// (no shared dt, converged lanes refilled).  Per lane it is the reference's
// DormandPrince45 (src/integrators.cpp:59-250) -- same tableau, FSAL, RMS error
// norm, controller and Hairer start -- with one algebraic rearrangement that
// halves the stage storage: because dx/dt = v, the stage velocities are
//   v_s = v + h * sum_l A[s][l] k_l            (k_l = stage accelerations)
// and the stage positions the reference builds from them collapse to
//   x_s = x + h * (c_s v + h * sum_l (A A)[s][l] k_l),
// while the position half of the embedded error is h^2 * sum_l (E A)[l] k_l
// (sum_s E_s = 0).  Only the seven k_l are stored.  This file is the CPU
// statement of that arithmetic, operation for operation, so the kernel can be
// checked tightly against it; the reference forms in scalar.cpp / batched.hpp
// pin it to the reference at the engine's accuracy bar.
//
// Also here: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11) and the engine's counter-based
// uniform-sphere sampler built on it (cube rejection + sequential
// min-separation placement, the placement rule of src/sampler.cpp:28-57).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "ce_oracle.h"
#include "oracle_common.hpp"

namespace ceo {
namespace {

struct LaneTableD {
  double c[7], aa[7][7], b[7], e[7], ea[7];
};

const LaneTableD& lane_table() {
  static const LaneTableD t = [] {
    const DP5Table& d = dp5_table();
    LaneTableD r{};
    const double c[7] = {0.0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
    for (int s = 0; s < 7; ++s) r.c[s] = c[s];
    // (A A)[s][l] = sum_{l<j<s} a[s][j] a[j][l] and (E A)[l] = sum_{s>l} e[s] a[s][l],
    // as exact rationals (tests/test_oracle.py recomputes them from the tableau in
    // rational arithmetic).
    const double aa[7][7] = {
        {0, 0, 0, 0, 0, 0, 0},
        {0, 0, 0, 0, 0, 0, 0},
        {9.0 / 200, 0, 0, 0, 0, 0, 0},
        {-12.0 / 25, 4.0 / 5, 0, 0, 0, 0, 0},
        {-12248.0 / 6561, 7208.0 / 2187, -6784.0 / 6561, 0, 0, 0, 0},
        {-533.0 / 264, 91.0 / 22, -56.0 / 33, 7.0 / 88, 0, 0, 0},
        {35.0 / 384, 0, 50.0 / 159, 25.0 / 192, -243.0 / 6784, 0, 0}};
    const double ea[7] = {611.0 / 230400,     0, -514.0 / 83475, 391.0 / 38400,
                          -4617.0 / 1356800, -11.0 / 3360, 0};
    for (int s = 0; s < 7; ++s)
      for (int l = 0; l < 7; ++l) r.aa[s][l] = aa[s][l];
    for (int l = 0; l < 7; ++l) {
      r.b[l] = d.a[6][l];
      r.e[l] = d.e[l];
      r.ea[l] = ea[l];
    }
    return r;
  }();
  return t;
}

template <class T>
struct LaneCoef {
  T c[7], aa[7][7], b[7], e[7], ea[7];
};

template <class T>
const LaneCoef<T>& lane_coef() {
  static const LaneCoef<T> k = [] {
    const LaneTableD& d = lane_table();
    LaneCoef<T> r{};
    for (int s = 0; s < 7; ++s) {
      r.c[s] = static_cast<T>(d.c[s]);
      r.b[s] = static_cast<T>(d.b[s]);
      r.e[s] = static_cast<T>(d.e[s]);
      r.ea[s] = static_cast<T>(d.ea[s]);
      for (int l = 0; l < 7; ++l) r.aa[s][l] = static_cast<T>(d.aa[s][l]);
    }
    return r;
  }();
  return k;
}

constexpr int kMaxN = 16;

template <class T>
struct LaneMol {
  int n;
  T pc[kMaxN][kMaxN];  // k q_i q_j narrowed from double (1 in the uniform form)
  T im[kMaxN];         // 1/m_i narrowed from double (c/m_i in the uniform form)
  T pe_scale = T(1);   // the summed PE times this is the PE (c in the uniform form)
  double pcd[kMaxN][kMaxN];
  double md[kMaxN];
};

// Per-sim charges, as the kernel builds them: kq_i = (T)(k q_i), pc_ij = kq_i * (T)q_j,
// and the FP64 diagnostic constant (double)kq_i * (double)q_j.
template <class T>
void set_lane_charges(LaneMol<T>& r, const ceo_molecule* m, const double* q) {
  for (int i = 0; i < m->n; ++i) {
    const T kq = static_cast<T>(m->coulomb_k * static_cast<double>(static_cast<float>(q[i])));
    for (int j = 0; j < m->n; ++j) {
      const T qj = static_cast<T>(static_cast<float>(q[j]));
      r.pc[i][j] = kq * qj;
      r.pcd[i][j] = static_cast<double>(kq) * static_cast<double>(qj);
    }
  }
}

// The kernel's molecule constants.  With `uniform_ok` (the f32 kernel without
// per-sim charges), a molecule whose pairs all have the same constant c (finite,
// non-zero; always so for two atoms) takes the kernel's uniform form: forces
// summed with pair constant 1, c folded into im = (T)(c / m_i), PE = c * sum.
template <class T>
LaneMol<T> lane_mol(const ceo_molecule* m, bool uniform_ok) {
  LaneMol<T> r{};
  r.n = m->n;
  for (int i = 0; i < m->n; ++i) {
    r.md[i] = m->masses[i];
    r.im[i] = static_cast<T>(1.0 / m->masses[i]);
    for (int j = 0; j < m->n; ++j) {
      r.pcd[i][j] = m->coulomb_k * m->charges[i] * m->charges[j];
      r.pc[i][j] = static_cast<T>(r.pcd[i][j]);
    }
  }
  bool uniform = uniform_ok && m->n >= 2;
  const double c = uniform ? r.pcd[0][1] : 0.0;
  uniform = uniform && c != 0.0 && std::isfinite(c);
  for (int i = 0; uniform && i < m->n; ++i)
    for (int j = i + 1; j < m->n; ++j) uniform = uniform && r.pcd[i][j] == c;
  if (uniform) {
    r.pe_scale = static_cast<T>(c);
    for (int i = 0; i < m->n; ++i) {
      r.im[i] = static_cast<T>(c / m->masses[i]);
      for (int j = 0; j < m->n; ++j) r.pc[i][j] = T(1);
    }
  }
  return r;
}

// Accelerations (and optionally PE) of one lane: x, a are [3n] (atom-major).
template <class T>
void lane_force(const LaneMol<T>& m, const T* x, T* a, T* pe, bool apply_im = true) {
  const int n = m.n;
  for (int q = 0; q < 3 * n; ++q) a[q] = T(0);
  T u = T(0);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const T rx = x[3 * i] - x[3 * j];
      const T ry = x[3 * i + 1] - x[3 * j + 1];
      const T rz = x[3 * i + 2] - x[3 * j + 2];
      const T d2 = std::fma(rx, rx, std::fma(ry, ry, rz * rz));
      const T inv = T(1) / std::sqrt(d2);
      const T inv3 = (inv * inv) * inv;
      const T s = m.pc[i][j] * inv3;
      if (pe) u = std::fma(m.pc[i][j], inv, u);
      a[3 * i] = std::fma(rx, s, a[3 * i]);
      a[3 * i + 1] = std::fma(ry, s, a[3 * i + 1]);
      a[3 * i + 2] = std::fma(rz, s, a[3 * i + 2]);
      a[3 * j] = std::fma(-rx, s, a[3 * j]);
      a[3 * j + 1] = std::fma(-ry, s, a[3 * j + 1]);
      a[3 * j + 2] = std::fma(-rz, s, a[3 * j + 2]);
    }
  if (apply_im)
    for (int i = 0; i < n; ++i) {
      a[3 * i] *= m.im[i];
      a[3 * i + 1] *= m.im[i];
      a[3 * i + 2] *= m.im[i];
    }
  if (pe) *pe = u * m.pe_scale;
}

template <class T>
T lane_pe(const LaneMol<T>& m, const T* x) {
  T u = T(0);
  for (int i = 0; i < m.n; ++i)
    for (int j = i + 1; j < m.n; ++j) {
      const T rx = x[3 * i] - x[3 * j];
      const T ry = x[3 * i + 1] - x[3 * j + 1];
      const T rz = x[3 * i + 2] - x[3 * j + 2];
      const T d2 = std::fma(rx, rx, std::fma(ry, ry, rz * rz));
      u = std::fma(m.pc[i][j], T(1) / std::sqrt(d2), u);
    }
  return u * m.pe_scale;
}

// Hairer / scipy initial step (src/integrators.cpp:197-250) for one lane, in the
// scalar reference's accumulation order; needs k0 = a(x0) and v.
template <class T>
T lane_hairer(const LaneMol<T>& m, const T* x, const T* v, const T* a0, T rtol, T atol,
              T max_dt) {
  const int n = m.n;
  T d0s = T(0), d1s = T(0);
  for (int q = 0; q < 3 * n; ++q) {
    const T sx = atol + rtol * std::abs(x[q]);
    const T sv = atol + rtol * std::abs(v[q]);
    T r = x[q] / sx;
    d0s += r * r;
    r = v[q] / sv;
    d0s += r * r;
    r = v[q] / sx;
    d1s += r * r;
    r = a0[q] / sv;
    d1s += r * r;
  }
  const T den = static_cast<T>(6 * n);
  const T d0 = std::sqrt(d0s / den);
  const T d1 = std::sqrt(d1s / den);
  const T h0 = (d0 < T(1e-5) || d1 < T(1e-5)) ? T(1e-6) : T(0.01) * d0 / d1;
  T xs[3 * kMaxN], a1[3 * kMaxN];
  for (int q = 0; q < 3 * n; ++q) xs[q] = x[q] + v[q] * h0;
  lane_force(m, xs, a1, static_cast<T*>(nullptr));
  T d2s = T(0);
  for (int q = 0; q < 3 * n; ++q) {
    const T sx = atol + rtol * std::abs(x[q]);
    const T sv = atol + rtol * std::abs(v[q]);
    T r = (a0[q] * h0) / sx;
    d2s += r * r;
    r = (a1[q] - a0[q]) / sv;
    d2s += r * r;
  }
  const T d2 = std::sqrt(d2s / den) / h0;
  const T h1 = (d1 <= T(1e-15) && d2 <= T(1e-15)) ? std::max(T(1e-6), h0 * T(1e-3))
                                                   : std::pow(T(0.01) / std::max(d1, d2), T(0.2));
  return std::min(std::min(T(100) * h0, h1), max_dt);
}

// One sim.  With t_cut > 0 (momentum-or-t_max rule) the run first stops at
// t_cut as a run with t_max = t_cut would; a sim stopped there by the time
// limit then continues to cfg->t_max from that state -- same x, v, E0, time
// and counts, the FSAL force, and the step size the controller proposed after
// the step that landed on t_cut (unclamped) -- as the engine's resume does.
template <class T>
void lane_run_one(const LaneMol<T>& m, const ceo_config* cfg, const double* p0, size_t sim,
                  ceo_results* out, double t_cut = 0.0) {
  const LaneCoef<T>& K = lane_coef<T>();
  const int n = m.n, n3 = 3 * m.n;
  const T rtol = static_cast<T>(cfg->rtol), atol = static_cast<T>(cfg->atol);
  const T max_dt = static_cast<T>(cfg->max_dt);
  T x[3 * kMaxN], v[3 * kMaxN], k[7][3 * kMaxN], xn[3 * kMaxN], vn[3 * kMaxN];
  for (int q = 0; q < n3; ++q) {
    x[q] = static_cast<T>(p0[q]);
    v[q] = T(0);
  }
  T pe0;
  // The stages hold forces F = m a; 1/m enters through per-atom step factors.
  lane_force(m, x, k[0], &pe0, false);
  const double e0 = static_cast<double>(pe0);
  const double pe_stop = cfg->pe_stop_fraction * e0;
  const T tol2 = static_cast<T>(cfg->momentum_tol * cfg->momentum_tol);
  uint32_t steps = 0, attempts = 0;
  int nrej = 0;
  bool rej_once = false;
  uint8_t status = 255;
  double t = 0.0;
  std::vector<double> trace;
  T h;
  if (cfg->stop_rule == CEO_STOP_PE_FRACTION && e0 <= pe_stop) {
    status = CEO_OK_PE;
    h = T(0);
  } else if (cfg->max_steps == 0) {
    status = CEO_NOT_CONVERGED;
    h = T(0);
  } else if (cfg->h0_mode == CEO_H0_FIXED) {
    h = static_cast<T>(cfg->h0);
  } else {
    T a0[3 * kMaxN];
    for (int q = 0; q < n3; ++q) a0[q] = k[0][q] * m.im[q / 3];
    h = lane_hairer(m, x, v, a0, rtol, atol, max_dt);
  }
  const T inv6n = T(1) / static_cast<T>(6 * n);
  bool cut_pending = t_cut > 0.0 && cfg->stop_rule == CEO_STOP_MOMENTUM_OR_TMAX;
  // A continuation from a stop starts from at least the proposal of the run's
  // last attempt that was shortened to land on the time limit (the engine's
  // saved step size): growing back from a step clamped to a sliver would take
  // several steps, each a sliver-sized momentum change.
  T h_short = T(0);
  while (status == 255) {
    const double t_lim = cut_pending ? t_cut : cfg->t_max;
    T cap = max_dt;
    bool by_time = false;
    if (cfg->stop_rule == CEO_STOP_MOMENTUM_OR_TMAX) {
      const double rem = t_lim - t;
      if (rem <= cfg->max_dt) {
        cap = static_cast<T>(rem);
        by_time = true;
      }
    }
    const T hh = std::min(h, cap);
    if (by_time && h > cap) h_short = h;  // proposal of the last shortened attempt
    const T hh2 = hh * hh;
    T g[kMaxN], gv[kMaxN];  // h^2 / m_i and h / m_i
    for (int i = 0; i < n; ++i) {
      g[i] = hh2 * m.im[i];
      gv[i] = hh * m.im[i];
    }
    ++attempts;
    // Stages 2..6: positions x + h c_s v + (h^2/m) sum_l aa[s][l] F_l, then F(x_s).
    for (int s = 1; s < 6; ++s) {
      const T hc = hh * K.c[s];
      for (int q = 0; q < n3; ++q) {
        T acc = T(0);
        bool first = true;
        for (int l = 0; l < s; ++l) {
          if (K.aa[s][l] == T(0)) continue;
          acc = first ? K.aa[s][l] * k[l][q] : std::fma(K.aa[s][l], k[l][q], acc);
          first = false;
        }
        const T base = std::fma(hc, v[q], x[q]);
        xn[q] = first ? base : std::fma(g[q / 3], acc, base);
      }
      lane_force(m, xn, k[s], static_cast<T*>(nullptr), false);
    }
    // Stage 7 (FSAL): the 5th-order solution itself.
    for (int q = 0; q < n3; ++q) {
      T ax = T(0), av = T(0);
      bool fx = true, fv = true;
      for (int l = 0; l < 6; ++l) {
        if (K.aa[6][l] != T(0)) {
          ax = fx ? K.aa[6][l] * k[l][q] : std::fma(K.aa[6][l], k[l][q], ax);
          fx = false;
        }
        if (K.b[l] != T(0)) {
          av = fv ? K.b[l] * k[l][q] : std::fma(K.b[l], k[l][q], av);
          fv = false;
        }
      }
      xn[q] = std::fma(g[q / 3], ax, std::fma(hh, v[q], x[q]));
      vn[q] = std::fma(gv[q / 3], av, v[q]);
    }
    lane_force(m, xn, k[6], static_cast<T*>(nullptr), false);
    // Embedded error, RMS over the 6n components (squared norm compared with 1).
    T ex_sum = T(0), ev_sum = T(0);
    for (int q = 0; q < n3; ++q) {
      T ex = T(0), ev = T(0);
      bool f1 = true, f2 = true;
      for (int l = 0; l < 7; ++l) {
        if (K.ea[l] != T(0)) {
          ex = f1 ? K.ea[l] * k[l][q] : std::fma(K.ea[l], k[l][q], ex);
          f1 = false;
        }
        if (K.e[l] != T(0)) {
          ev = f2 ? K.e[l] * k[l][q] : std::fma(K.e[l], k[l][q], ev);
          f2 = false;
        }
      }
      const T scx = std::fma(rtol, std::max(std::abs(x[q]), std::abs(xn[q])), atol);
      const T scv = std::fma(rtol, std::max(std::abs(v[q]), std::abs(vn[q])), atol);
      const T rx = (g[q / 3] * ex) * (T(1) / scx);
      const T rv = (gv[q / 3] * ev) * (T(1) / scv);
      ex_sum = std::fma(rx, rx, ex_sum);
      ev_sum = std::fma(rv, rv, ev_sum);
    }
    const T en2 = (ex_sum + ev_sum) * inv6n;
    // 0.9 err^(-1/5) with err = sqrt(en2): 0.9 * 2^(-0.1 log2 en2).
    const T grow = T(0.9) * std::exp2(T(-0.1) * std::log2(en2));
    // The step that lands on the time limit (shortened to the remaining time, or
    // ending there in FP64) is final; the momentum test is not applied to it.
    const bool final_step = by_time && (hh >= cap || t + static_cast<double>(hh) >= t_lim);
    if (en2 <= T(1)) {
      bool settled = true;
      if (cfg->stop_rule == CEO_STOP_MOMENTUM_OR_TMAX) {
        settled = !final_step;
        for (int i = 0; i < n && settled; ++i) {
          const T dx = vn[3 * i] - v[3 * i], dy = vn[3 * i + 1] - v[3 * i + 1],
                  dz = vn[3 * i + 2] - v[3 * i + 2];
          const T num = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
          const T den = std::fma(vn[3 * i + 2], vn[3 * i + 2],
                                 std::fma(vn[3 * i + 1], vn[3 * i + 1], vn[3 * i] * vn[3 * i]));
          settled = num < tol2 * den;
        }
      }
      for (int q = 0; q < n3; ++q) {
        x[q] = xn[q];
        v[q] = vn[q];
        k[0][q] = k[6][q];
      }
      T f = (en2 == T(0)) ? T(10) : std::min(T(10), grow);
      if (rej_once) f = std::min(f, T(1));
      // After the step that lands on the time limit the step size a continuation
      // starts with is not clamped to the limit.
      h = std::min(hh * f, final_step ? max_dt : cap);
      rej_once = false;
      nrej = 0;
      ++steps;
      trace.push_back(static_cast<double>(hh));
      t = final_step ? t_lim : t + static_cast<double>(hh);
      if (cfg->stop_rule == CEO_STOP_PE_FRACTION) {
        if (static_cast<double>(lane_pe(m, x)) <= pe_stop) status = CEO_OK_PE;
      } else if (settled) {
        status = CEO_OK_MOMENTUM;
      } else if (final_step && cut_pending) {
        cut_pending = false;  // continue from the cut to cfg->t_max
        h = std::max(h, h_short);
        h_short = T(0);
      } else if (final_step) {
        status = CEO_OK_TMAX;
      }
      if (status == 255 && steps >= cfg->max_steps) status = CEO_NOT_CONVERGED;
    } else {
      const T f = std::max(T(0.2), grow);
      h = hh * f;
      rej_once = true;
      if (++nrej > cfg->max_rejections) status = std::isnan(en2) ? CEO_NONFINITE : CEO_REJECT_LIMIT;
    }
  }
  // Finalize in double: redistribution s = sqrt(E0/KE), p = m s v.
  double ke = 0.0;
  for (int i = 0; i < n; ++i) {
    const double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
    ke += 0.5 * m.md[i] * (vx * vx + vy * vy + vz * vz);
  }
  double s = 1.0;
  if (cfg->redistribute && ke > 0.0) s = std::sqrt(e0 / ke);
  double pe_end = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double rx = double(x[3 * i]) - double(x[3 * j]);
      const double ry = double(x[3 * i + 1]) - double(x[3 * j + 1]);
      const double rz = double(x[3 * i + 2]) - double(x[3 * j + 2]);
      pe_end += m.pcd[i][j] / std::sqrt(rx * rx + ry * ry + rz * rz);
    }
  for (int i = 0; i < n; ++i) {
    const double ms = m.md[i] * s;
    double* p = out->momenta + (sim * n + i) * 3;
    p[0] = ms * double(v[3 * i]);
    p[1] = ms * double(v[3 * i + 1]);
    p[2] = ms * double(v[3 * i + 2]);
  }
  out->steps[sim] = steps;
  out->status[sim] = status;
  if (out->attempts) out->attempts[sim] = attempts;
  if (out->t_final) out->t_final[sim] = t;
  if (out->energy_err) out->energy_err[sim] = (ke + pe_end - e0) / e0;
  if (out->trace_h)
    for (uint32_t q = 0; q < out->trace_cap; ++q)
      out->trace_h[sim * out->trace_cap + q] = q < trace.size() ? trace[q] : 0.0;
}

template <class T>
int lane_run(const ceo_molecule* cm, const ceo_config* cfg, const double* pos, size_t n_sims,
             ceo_results* out, const double* charges, double t_cut = 0.0) {
  if (cm->n < 1 || cm->n > kMaxN) return -1;
  LaneMol<T> m = lane_mol<T>(cm, std::is_same<T, float>::value && charges == nullptr);
  for (size_t sim = 0; sim < n_sims; ++sim) {
    if (charges) set_lane_charges<T>(m, cm, charges + sim * cm->n);
    lane_run_one<T>(m, cfg, pos + sim * cm->n * 3, sim, out, t_cut);
  }
  return 0;
}

inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
  const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
  const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
  const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

}  // namespace
}  // namespace ceo

using namespace ceo;

extern "C" {

int ceo_lane_run_f32(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                     size_t n_sims, ceo_results* out) {
  return lane_run<float>(mol, cfg, pos, n_sims, out, nullptr);
}

int ceo_lane_run_f64(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                     size_t n_sims, ceo_results* out) {
  return lane_run<double>(mol, cfg, pos, n_sims, out, nullptr);
}

int ceo_lane_run_q_f32(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                       const double* charges, size_t n_sims, ceo_results* out) {
  return lane_run<float>(mol, cfg, pos, n_sims, out, charges);
}

int ceo_lane_run_q_f64(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                       const double* charges, size_t n_sims, ceo_results* out) {
  return lane_run<double>(mol, cfg, pos, n_sims, out, charges);
}

int ceo_lane_run_cut(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                     const double* charges, size_t n_sims, double t_cut, int32_t precision,
                     ceo_results* out) {
  if (!(t_cut > 0.0) || cfg->stop_rule != CEO_STOP_MOMENTUM_OR_TMAX) return -1;
  return precision == 1 ? lane_run<double>(mol, cfg, pos, n_sims, out, charges, t_cut)
                        : lane_run<float>(mol, cfg, pos, n_sims, out, charges, t_cut);
}

void ceo_lane_coefficients(double* c, double* aa, double* b, double* e, double* ea) {
  const LaneTableD& t = lane_table();
  for (int s = 0; s < 7; ++s) {
    c[s] = t.c[s];
    b[s] = t.b[s];
    e[s] = t.e[s];
    ea[s] = t.ea[s];
    for (int l = 0; l < 7; ++l) aa[s * 7 + l] = t.aa[s][l];
  }
}

void ceo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    philox_round(c, k);
  }
  std::memcpy(out, c, sizeof(c));
}

// The engine's counter-based uniform-sphere sampler.  Per simulation g, candidate
// number j uses Philox4x32-10 with counter (j, lo32(g), hi32(g), 0) and key
// (lo32(seed), hi32(seed)); words 0..2 -> 24-bit uniforms u in [0,1) ->
// coordinates fma(u, 2R, -R) in float.  A candidate is kept if it lies in the ball
// (cube rejection) and clears every atom placed before it by min_sep
// (src/sampler.cpp:28-57); an atom failing max_attempts separation tests, or a
// candidate needing more than 4096 cube draws, marks the sim CEO_SAMPLER_FAIL.
size_t ceo_device_sample_q(uint64_t seed, uint64_t first_index, size_t n_sims, int32_t n_atoms,
                           double radius, double min_sep, uint32_t max_attempts,
                           int32_t charge_levels, int32_t max_total, double* pos_out,
                           double* q_out, uint8_t* status_out);

size_t ceo_device_sample(uint64_t seed, uint64_t first_index, size_t n_sims, int32_t n_atoms,
                         double radius, double min_sep, uint32_t max_attempts, double* pos_out,
                         uint8_t* status_out) {
  return ceo_device_sample_q(seed, first_index, n_sims, n_atoms, radius, min_sep, max_attempts,
                             0, 0, pos_out, nullptr, status_out);
}

// As ceo_device_sample, plus per-sim charges when q_out is given: try t draws
// Philox4x32-10 on counters (4t + i/4, lo32(g), hi32(g), 1), word i % 4, and
// q_i = 1 + ((u24 * levels) >> 24); tries continue while the sum exceeds
// max_total (> 0), at most 4096.
size_t ceo_device_sample_q(uint64_t seed, uint64_t first_index, size_t n_sims, int32_t n_atoms,
                           double radius, double min_sep, uint32_t max_attempts,
                           int32_t charge_levels, int32_t max_total, double* pos_out,
                           double* q_out, uint8_t* status_out) {
  const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
  const float R = static_cast<float>(radius);
  const float twoR = static_cast<float>(2.0 * radius);
  const float R2 = R * R;
  const float ms = static_cast<float>(min_sep);
  const float min2 = ms * ms;
  size_t fails = 0;
  for (size_t s = 0; s < n_sims; ++s) {
    const uint64_t g = first_index + s;
    uint32_t cand = 0;
    float P[3 * kMaxN];
    bool ok_sim = true;
    for (int i = 0; i < n_atoms && ok_sim; ++i) {
      uint32_t tries = 0;
      for (;;) {
        float c[3];
        bool in_ball = false;
        for (int cube = 0; cube < 4096; ++cube) {
          const uint32_t ctr[4] = {cand++, static_cast<uint32_t>(g), static_cast<uint32_t>(g >> 32),
                                   0u};
          uint32_t r[4];
          ceo_philox4x32_10(ctr, key, r);
          for (int q = 0; q < 3; ++q) {
            const float u = static_cast<float>(r[q] >> 8) * 0x1.0p-24f;
            c[q] = std::fma(u, twoR, -R);
          }
          const float d = std::fma(c[2], c[2], std::fma(c[1], c[1], c[0] * c[0]));
          if (d <= R2) {
            in_ball = true;
            break;
          }
        }
        if (!in_ball) {
          ok_sim = false;
          break;
        }
        bool clears = true;
        for (int j = 0; j < i && clears; ++j) {
          const float dx = c[0] - P[3 * j], dy = c[1] - P[3 * j + 1], dz = c[2] - P[3 * j + 2];
          const float d2 = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
          clears = !(d2 < min2);
        }
        if (clears) {
          P[3 * i] = c[0];
          P[3 * i + 1] = c[1];
          P[3 * i + 2] = c[2];
          break;
        }
        if (++tries >= max_attempts) {
          ok_sim = false;
          break;
        }
      }
    }
    if (q_out) {
      double Q[kMaxN];
      bool placed = false;
      for (uint32_t t = 0; t < 4096 && !placed; ++t) {
        int total = 0;
        for (int b = 0; b < (n_atoms + 3) / 4; ++b) {
          const uint32_t ctr[4] = {4u * t + static_cast<uint32_t>(b), static_cast<uint32_t>(g),
                                   static_cast<uint32_t>(g >> 32), 1u};
          uint32_t r[4];
          ceo_philox4x32_10(ctr, key, r);
          for (int w = 0; w < 4; ++w) {
            const int i = 4 * b + w;
            if (i >= n_atoms) break;
            const int qi = 1 + static_cast<int>((static_cast<uint64_t>(r[w] >> 8) *
                                                 static_cast<uint32_t>(charge_levels)) >> 24);
            Q[i] = qi;
            total += qi;
          }
        }
        placed = max_total <= 0 || total <= max_total;
      }
      ok_sim = ok_sim && placed;
      for (int i = 0; i < n_atoms; ++i) q_out[s * n_atoms + i] = ok_sim ? Q[i] : 0.0;
    }
    double* out = pos_out + s * n_atoms * 3;
    for (int q = 0; q < 3 * n_atoms; ++q) out[q] = ok_sim ? static_cast<double>(P[q]) : 0.0;
    if (status_out) status_out[s] = ok_sim ? 0 : CEO_SAMPLER_FAIL;
    if (!ok_sim) ++fails;
  }
  return fails;
}

}  // extern "C"
