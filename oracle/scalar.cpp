// CPU oracle, scalar engine -- TEST INFRASTRUCTURE ONLY (see ce_oracle.h).
//
// Restates, operation for operation, the reference's scalar path so that a build
// of this file and a build of the reference sources with the same compiler flags
// give bit-identical results:
//   force / PE / KE ............ src/system.cpp:5-49
//   DormandPrince45 ............ src/integrators.cpp:59-250
//   run_to_convergence ......... src/driver.cpp:7-46
//   UniformSphereSampler ....... src/sampler.cpp:7-57
// Precision follows the reference's compile-time `Real`: T = double is the
// default build, T = float the COULOMB_SINGLE_PRECISION build
// (include/coulomb/types.hpp:16-20).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#include "ce_oracle.h"
#include "oracle_common.hpp"

namespace ceo {
namespace {

// One atom's 3-vector in the reference's component order.
template <class T>
struct V3 {
  T x{0}, y{0}, z{0};
};
template <class T>
inline T dot3(const V3<T>& a, const V3<T>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;  // (xx + yy) + zz, as types.hpp:49
}

template <class T>
struct Mol {
  int n;
  std::vector<T> mass, charge;
  T k;
};

template <class T>
Mol<T> make_mol(const ceo_molecule* m) {
  Mol<T> r;
  r.n = m->n;
  r.k = static_cast<T>(m->coulomb_k);
  for (int i = 0; i < m->n; ++i) {
    r.mass.push_back(static_cast<T>(m->masses[i]));
    r.charge.push_back(static_cast<T>(m->charges[i]));
  }
  return r;
}

// system.cpp:5-29 -- acceleration of every atom, Newton-3 over i<j.
template <class T>
void accel(const Mol<T>& m, const std::vector<V3<T>>& x, std::vector<V3<T>>& a) {
  const int n = m.n;
  a.assign(n, V3<T>{});
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      V3<T> r{x[i].x - x[j].x, x[i].y - x[j].y, x[i].z - x[j].z};
      const T d2 = dot3(r, r);
      const T inv = T(1) / std::sqrt(d2);
      const T inv3 = inv / d2;
      const T s = m.k * m.charge[i] * m.charge[j] * inv3;
      const T fx = r.x * s, fy = r.y * s, fz = r.z * s;
      a[i].x += fx;
      a[i].y += fy;
      a[i].z += fz;
      a[j].x -= fx;
      a[j].y -= fy;
      a[j].z -= fz;
    }
  }
  for (int i = 0; i < n; ++i) {
    const T im = T(1) / m.mass[i];
    a[i].x = a[i].x * im;
    a[i].y = a[i].y * im;
    a[i].z = a[i].z * im;
  }
}

// system.cpp:31-41
template <class T>
T potential(const Mol<T>& m, const std::vector<V3<T>>& x) {
  T u{0};
  for (int i = 0; i < m.n; ++i)
    for (int j = i + 1; j < m.n; ++j) {
      V3<T> r{x[i].x - x[j].x, x[i].y - x[j].y, x[i].z - x[j].z};
      u += m.k * m.charge[i] * m.charge[j] / std::sqrt(dot3(r, r));
    }
  return u;
}

// system.cpp:43-49
template <class T>
T kinetic(const Mol<T>& m, const std::vector<V3<T>>& v) {
  T t{0};
  for (int i = 0; i < m.n; ++i) t += T(0.5) * m.mass[i] * dot3(v[i], v[i]);
  return t;
}

// integrators.cpp:59-250 -- adaptive DP5(4), scipy-compatible controller.
template <class T>
class ScalarDP45 {
 public:
  ScalarDP45(T rtol, T atol, int max_rej) : rtol_(rtol), atol_(atol), max_rej_(max_rej) {}

  void set_fixed_h0(T h0) { fixed_h0_ = h0; }

  // Returns the accepted step, or a negative status on failure.
  T step(const Mol<T>& m, std::vector<V3<T>>& x, std::vector<V3<T>>& v, T max_dt, int* fail,
         uint32_t* attempts) {
    const int n = m.n;
    if (static_cast<int>(kp_[0].size()) != n) {
      for (auto& s : kp_) s.assign(n, V3<T>{});
      for (auto& s : kv_) s.assign(n, V3<T>{});
      stage_.assign(n, V3<T>{});
      fsal_ = false;
      h_ = T(0);
    }
    if (!fsal_) {
      kp_[0] = v;
      accel(m, x, kv_[0]);
      fsal_ = true;
    }
    if (h_ <= T(0)) h_ = (fixed_h0_ > T(0)) ? fixed_h0_ : initial_step(m, x, v, max_dt);

    bool rejected_once = false;
    for (int attempt = 0;; ++attempt) {
      if (attempt > max_rej_) {
        *fail = 1;
        return T(-1);
      }
      const T h = std::min(h_, max_dt);
      if (attempts) ++*attempts;
      for (int s = 1; s < 7; ++s) {
        const T* a = coef<T>().a[s];
        for (int i = 0; i < n; ++i) {
          V3<T> dx{}, dv{};
          for (int j = 0; j < s; ++j) {
            const T c = a[j];
            dx.x += kp_[j][i].x * c;
            dx.y += kp_[j][i].y * c;
            dx.z += kp_[j][i].z * c;
            dv.x += kv_[j][i].x * c;
            dv.y += kv_[j][i].y * c;
            dv.z += kv_[j][i].z * c;
          }
          stage_[i] = V3<T>{x[i].x + dx.x * h, x[i].y + dx.y * h, x[i].z + dx.z * h};
          kp_[s][i] = V3<T>{v[i].x + dv.x * h, v[i].y + dv.y * h, v[i].z + dv.z * h};
        }
        accel(m, stage_, kv_[s]);
      }
      T err_sq{0};
      for (int i = 0; i < n; ++i) {
        V3<T> ep{}, ev{};
        for (int s = 0; s < 7; ++s) {
          const T c = coef<T>().e[s];
          ep.x += kp_[s][i].x * c;
          ep.y += kp_[s][i].y * c;
          ep.z += kp_[s][i].z * c;
          ev.x += kv_[s][i].x * c;
          ev.y += kv_[s][i].y * c;
          ev.z += kv_[s][i].z * c;
        }
        err_sq += errc(ep.x * h, x[i].x, stage_[i].x);
        err_sq += errc(ep.y * h, x[i].y, stage_[i].y);
        err_sq += errc(ep.z * h, x[i].z, stage_[i].z);
        err_sq += errc(ev.x * h, v[i].x, kp_[6][i].x);
        err_sq += errc(ev.y * h, v[i].y, kp_[6][i].y);
        err_sq += errc(ev.z * h, v[i].z, kp_[6][i].z);
      }
      const T en = std::sqrt(err_sq / static_cast<T>(6 * n));
      if (en <= T(1)) {
        x.swap(stage_);
        v.swap(kp_[6]);
        kv_[0].swap(kv_[6]);
        kp_[0] = v;
        T f = (en == T(0)) ? T(10) : std::min(T(10), T(0.9) * std::pow(en, T(-1.0 / 5.0)));
        if (rejected_once) f = std::min(f, T(1));
        h_ = std::min(h * f, max_dt);
        return h;
      }
      last_err_nan_ = std::isnan(en);
      const T f = std::max(T(0.2), T(0.9) * std::pow(en, T(-1.0 / 5.0)));
      h_ = h * f;
      rejected_once = true;
    }
  }
  bool last_err_nan() const { return last_err_nan_; }

 private:
  T errc(T err, T yo, T yn) const {
    const T sc = atol_ + rtol_ * std::max(std::abs(yo), std::abs(yn));
    const T r = err / sc;
    return r * r;
  }

  // integrators.cpp:197-250 (Hairer / scipy select_initial_step).
  T initial_step(const Mol<T>& m, const std::vector<V3<T>>& x, const std::vector<V3<T>>& v,
                 T max_dt) {
    const int n = m.n;
    T d0s{0}, d1s{0};
    auto acc = [](T& s, T val, T sc) {
      const T r = val / sc;
      s += r * r;
    };
    for (int i = 0; i < n; ++i) {
      const T xc[3] = {x[i].x, x[i].y, x[i].z};
      const T vc[3] = {v[i].x, v[i].y, v[i].z};
      const T fp[3] = {kp_[0][i].x, kp_[0][i].y, kp_[0][i].z};
      const T fv[3] = {kv_[0][i].x, kv_[0][i].y, kv_[0][i].z};
      for (int c = 0; c < 3; ++c) {
        const T sx = atol_ + rtol_ * std::abs(xc[c]);
        const T sv = atol_ + rtol_ * std::abs(vc[c]);
        acc(d0s, xc[c], sx);
        acc(d0s, vc[c], sv);
        acc(d1s, fp[c], sx);
        acc(d1s, fv[c], sv);
      }
    }
    const T den = static_cast<T>(6 * n);
    const T d0 = std::sqrt(d0s / den);
    const T d1 = std::sqrt(d1s / den);
    const T h0 = (d0 < 1e-5 || d1 < 1e-5) ? T(1e-6) : T(0.01) * d0 / d1;
    for (int i = 0; i < n; ++i)
      stage_[i] = V3<T>{x[i].x + kp_[0][i].x * h0, x[i].y + kp_[0][i].y * h0,
                        x[i].z + kp_[0][i].z * h0};
    std::vector<V3<T>> a1;
    accel(m, stage_, a1);
    T d2s{0};
    for (int i = 0; i < n; ++i) {
      const T dvp[3] = {kv_[0][i].x * h0, kv_[0][i].y * h0, kv_[0][i].z * h0};
      const T dap[3] = {a1[i].x - kv_[0][i].x, a1[i].y - kv_[0][i].y, a1[i].z - kv_[0][i].z};
      const T xc[3] = {x[i].x, x[i].y, x[i].z};
      const T vc[3] = {v[i].x, v[i].y, v[i].z};
      for (int c = 0; c < 3; ++c) {
        const T sx = atol_ + rtol_ * std::abs(xc[c]);
        const T sv = atol_ + rtol_ * std::abs(vc[c]);
        acc(d2s, dvp[c], sx);
        acc(d2s, dap[c], sv);
      }
    }
    const T d2 = std::sqrt(d2s / den) / h0;
    const T h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? std::max(T(1e-6), h0 * T(1e-3))
                                               : std::pow(T(0.01) / std::max(d1, d2), T(1.0 / 5.0));
    return std::min({100 * h0, h1, max_dt});
  }

  T rtol_, atol_;
  int max_rej_;
  T h_{0};
  T fixed_h0_{0};
  bool fsal_{false};
  bool last_err_nan_{false};
  std::vector<V3<T>> kp_[7], kv_[7], stage_;
};

// driver.cpp:7-46, with the t_max / momentum stop rule and FP64 time added.
template <class T>
int scalar_run(const ceo_molecule* cm, const ceo_config* cfg, const double* pos, size_t n_sims,
               ceo_results* out) {
  const Mol<T> m = make_mol<T>(cm);
  const int n = m.n;
  for (size_t sim = 0; sim < n_sims; ++sim) {
    std::vector<V3<T>> x(n), v(n), vold(n);
    for (int i = 0; i < n; ++i) {
      const double* p = pos + (sim * n + i) * 3;
      x[i] = V3<T>{static_cast<T>(p[0]), static_cast<T>(p[1]), static_cast<T>(p[2])};
    }
    ScalarDP45<T> integ(static_cast<T>(cfg->rtol), static_cast<T>(cfg->atol),
                        cfg->max_rejections);
    if (cfg->h0_mode == CEO_H0_FIXED) integ.set_fixed_h0(static_cast<T>(cfg->h0));
    const T pe0 = potential(m, x);
    const T e0 = pe0 + kinetic(m, v);
    uint32_t steps = 0, attempts = 0;
    uint8_t status = CEO_NOT_CONVERGED;
    double t = 0.0;
    std::vector<double> trace;
    if (cfg->stop_rule == CEO_STOP_PE_FRACTION) {
      const T pe_stop = static_cast<T>(cfg->pe_stop_fraction) * pe0;
      T pe = pe0;
      T tf{0};
      bool conv = false;
      while (pe > pe_stop) {
        if (steps >= cfg->max_steps) {
          conv = false;
          break;
        }
        int fail = 0;
        const T dt = integ.step(m, x, v, static_cast<T>(cfg->max_dt), &fail, &attempts);
        if (fail) {
          status = integ.last_err_nan() ? CEO_NONFINITE : CEO_REJECT_LIMIT;
          break;
        }
        trace.push_back(dt);
        tf += dt;
        ++steps;
        pe = potential(m, x);
        conv = true;
      }
      if (status != CEO_REJECT_LIMIT && status != CEO_NONFINITE)
        status = conv ? CEO_OK_PE : CEO_NOT_CONVERGED;
      // The scalar driver leaves converged=false when PE0 <= pe_stop already
      // (driver.cpp:19-30); report that case as CEO_OK_PE like the batched engine.
      if (steps == 0 && !(pe0 > pe_stop)) status = CEO_OK_PE;
      t = static_cast<double>(tf);
    } else {
      const T tol2 = static_cast<T>(cfg->momentum_tol * cfg->momentum_tol);
      for (;;) {
        if (steps >= cfg->max_steps) {
          status = CEO_NOT_CONVERGED;
          break;
        }
        const double rem = cfg->t_max - t;
        const bool cap_by_time = rem <= cfg->max_dt;
        const T cap = static_cast<T>(cap_by_time ? rem : cfg->max_dt);
        vold = v;
        int fail = 0;
        const T dt = integ.step(m, x, v, cap, &fail, &attempts);
        if (fail) {
          status = integ.last_err_nan() ? CEO_NONFINITE : CEO_REJECT_LIMIT;
          break;
        }
        trace.push_back(dt);
        ++steps;
        const bool final_step =
            cap_by_time && (dt >= cap || t + static_cast<double>(dt) >= cfg->t_max);
        t = final_step ? cfg->t_max : t + static_cast<double>(dt);
        // The momentum test is not applied to the step that lands on t_max.
        if (!final_step && momentum_settled(n, vold, v, tol2)) {
          status = CEO_OK_MOMENTUM;
          break;
        }
        if (final_step) {
          status = CEO_OK_TMAX;
          break;
        }
      }
    }
    // Finalize (driver.cpp:32-43) in the engine's precision.
    const T pe_end = potential(m, x);
    const T ke = kinetic(m, v);
    if (cfg->redistribute && ke > T(0)) {
      const T s = std::sqrt(e0 / ke);
      for (auto& w : v) w = V3<T>{w.x * s, w.y * s, w.z * s};
    }
    for (int i = 0; i < n; ++i) {
      double* p = out->momenta + (sim * n + i) * 3;
      p[0] = static_cast<double>(v[i].x * m.mass[i]);
      p[1] = static_cast<double>(v[i].y * m.mass[i]);
      p[2] = static_cast<double>(v[i].z * m.mass[i]);
    }
    out->steps[sim] = steps;
    out->status[sim] = status;
    if (out->attempts) out->attempts[sim] = attempts;
    if (out->t_final) out->t_final[sim] = t;
    if (out->energy_err)
      out->energy_err[sim] = (static_cast<double>(ke) + static_cast<double>(pe_end) -
                              static_cast<double>(e0)) /
                             static_cast<double>(e0);
    if (out->trace_h)
      for (uint32_t k = 0; k < out->trace_cap; ++k)
        out->trace_h[sim * out->trace_cap + k] = k < trace.size() ? trace[k] : 0.0;
  }
  return 0;
}

}  // namespace
}  // namespace ceo

using namespace ceo;

extern "C" {

void ceo_accelerations_f64(const ceo_molecule* cm, const double* pos, double* acc) {
  const Mol<double> m = make_mol<double>(cm);
  std::vector<V3<double>> x(m.n), a;
  for (int i = 0; i < m.n; ++i) x[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  accel(m, x, a);
  for (int i = 0; i < m.n; ++i) {
    acc[3 * i] = a[i].x;
    acc[3 * i + 1] = a[i].y;
    acc[3 * i + 2] = a[i].z;
  }
}

void ceo_accelerations_f32(const ceo_molecule* cm, const float* pos, float* acc) {
  const Mol<float> m = make_mol<float>(cm);
  std::vector<V3<float>> x(m.n), a;
  for (int i = 0; i < m.n; ++i) x[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  accel(m, x, a);
  for (int i = 0; i < m.n; ++i) {
    acc[3 * i] = a[i].x;
    acc[3 * i + 1] = a[i].y;
    acc[3 * i + 2] = a[i].z;
  }
}

double ceo_potential_energy_f64(const ceo_molecule* cm, const double* pos) {
  const Mol<double> m = make_mol<double>(cm);
  std::vector<V3<double>> x(m.n);
  for (int i = 0; i < m.n; ++i) x[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  return potential(m, x);
}

float ceo_potential_energy_f32(const ceo_molecule* cm, const float* pos) {
  const Mol<float> m = make_mol<float>(cm);
  std::vector<V3<float>> x(m.n);
  for (int i = 0; i < m.n; ++i) x[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  return potential(m, x);
}

double ceo_kinetic_energy_f64(const ceo_molecule* cm, const double* vel) {
  const Mol<double> m = make_mol<double>(cm);
  std::vector<V3<double>> v(m.n);
  for (int i = 0; i < m.n; ++i) v[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
  return kinetic(m, v);
}

// sampler.cpp:7-57.  The stream is std::mt19937_64 feeding
// std::uniform_real_distribution<double>(-R, R), three draws per candidate.
int ceo_sample_uniform_sphere(uint64_t seed, int32_t n_atoms, double radius, double min_sep,
                              uint64_t max_attempts, size_t n_sims, double* pos_out) {
  if (radius <= 0.0 || min_sep < 0.0) return -1;
  std::mt19937_64 rng(seed);
  const double r2 = radius * radius;
  const double min2 = min_sep * min_sep;
  for (size_t sim = 0; sim < n_sims; ++sim) {
    double* P = pos_out + sim * n_atoms * 3;
    for (int i = 0; i < n_atoms; ++i) {
      uint64_t tries = 0;
      for (;;) {
        double c[3];
        for (;;) {
          std::uniform_real_distribution<double> u(-radius, radius);
          c[0] = u(rng);
          c[1] = u(rng);
          c[2] = u(rng);
          if (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] <= r2) break;
        }
        bool ok = true;
        for (int j = 0; j < i && ok; ++j) {
          const double dx = c[0] - P[3 * j], dy = c[1] - P[3 * j + 1], dz = c[2] - P[3 * j + 2];
          if (dx * dx + dy * dy + dz * dz < min2) ok = false;
        }
        if (ok) {
          P[3 * i] = c[0];
          P[3 * i + 1] = c[1];
          P[3 * i + 2] = c[2];
          break;
        }
        if (++tries >= max_attempts) return -2;
      }
    }
  }
  return 0;
}

int ceo_scalar_run_f64(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                       size_t n_sims, ceo_results* out) {
  return scalar_run<double>(mol, cfg, pos, n_sims, out);
}

int ceo_scalar_run_f32(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                       size_t n_sims, ceo_results* out) {
  return scalar_run<float>(mol, cfg, pos, n_sims, out);
}

}  // extern "C"
