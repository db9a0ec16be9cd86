// CPU oracle, lockstep batched engine -- TEST INFRASTRUCTURE ONLY (see ce_oracle.h).
//
// Restates BatchedRK45<T> (include/coulomb/batched_integrator.hpp:62-433) on the
// batched kernels of include/coulomb/batched_force.hpp:31-110 without Highway:
// K lanes, one simulation per lane, SoA p[i*K + lane], every lane loop innermost
// so the compiler vectorises it at the width the translation unit targets.
// Arithmetic per lane follows the reference's Highway calls: MulAdd -> std::fma,
// Div(1, Sqrt) -> 1/std::sqrt, the controller's pow in double.
//
// Added to the reference semantics (documented extensions, used by the
// engine's headline configuration):
//   * stop rule CEO_STOP_MOMENTUM_OR_TMAX with one FP64 batch time (all lanes
//     share dt, so they share t);
//   * a fixed initial step (CEO_H0_FIXED) instead of the Hairer heuristic.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "ce_oracle.h"
#include "oracle_common.hpp"

namespace ceo {

template <class T, int K>
class Lockstep {
 public:
  struct Soa3 {
    std::vector<T> x, y, z;
    void alloc(size_t m) {
      x.assign(m, T(0));
      y.assign(m, T(0));
      z.assign(m, T(0));
    }
    void swap(Soa3& o) {
      x.swap(o.x);
      y.swap(o.y);
      z.swap(o.z);
    }
    T* c(int i) { return i == 0 ? x.data() : (i == 1 ? y.data() : z.data()); }
  };

  Lockstep(const ceo_molecule* m, T rtol, T atol, int max_rej)
      : n_(m->n), rtol_(rtol), atol_(atol), max_rej_(max_rej) {
    pc_.assign(n_ * n_, T(0));
    im_.assign(n_, T(0));
    md_.assign(n_, 0.0);
    for (int i = 0; i < n_; ++i) {
      md_[i] = m->masses[i];
      im_[i] = static_cast<T>(1.0 / m->masses[i]);
      for (int j = 0; j < n_; ++j)
        pc_[i * n_ + j] = static_cast<T>(m->coulomb_k * m->charges[i] * m->charges[j]);
    }
    const size_t sz = static_cast<size_t>(n_) * K;
    pos_.alloc(sz);
    vel_.alloc(sz);
    stage_.alloc(sz);
    probe_.alloc(sz);
    fin_.alloc(sz);
    vold_.alloc(sz);
    for (int s = 0; s < 7; ++s) {
      kp_[s].alloc(sz);
      kv_[s].alloc(sz);
    }
  }

  void set_lane(int lane, const double* p) {
    for (int i = 0; i < n_; ++i) {
      pos_.x[i * K + lane] = static_cast<T>(p[3 * i]);
      pos_.y[i * K + lane] = static_cast<T>(p[3 * i + 1]);
      pos_.z[i * K + lane] = static_cast<T>(p[3 * i + 2]);
      vel_.x[i * K + lane] = T(0);
      vel_.y[i * K + lane] = T(0);
      vel_.z[i * K + lane] = T(0);
    }
  }

  struct LaneOut {
    double mom[3 * 16];
    uint32_t steps;
    uint8_t status;
    double t;
    double eerr;
  };

  // One batch to its stop.  Returns the number of accepted batch iterations,
  // or -1 on a rejection storm (the reference throws runtime_error).
  int64_t run(const ceo_config* cfg, LaneOut* out, uint32_t* attempts_total) {
    T pe[K];
    potential(pos_, pe);
    bool done[K];
    int n_done = 0;
    for (int l = 0; l < K; ++l) {
      pe0_[l] = static_cast<double>(pe[l]);
      pes_[l] = cfg->pe_stop_fraction * pe0_[l];
      e0_[l] = pe0_[l];
      done[l] = false;
      out[l].status = CEO_NOT_CONVERGED;
      out[l].steps = 0;
      out[l].t = 0.0;
      if (cfg->stop_rule == CEO_STOP_PE_FRACTION && pe0_[l] <= pes_[l]) {
        done[l] = true;
        out[l].status = CEO_OK_PE;
        latch(l);
        ++n_done;
      }
    }
    h_ = T(0);
    fsal_ = false;
    uint64_t iter = 0;
    double t = 0.0;
    const T tol2 = static_cast<T>(cfg->momentum_tol * cfg->momentum_tol);
    while (n_done < K && iter < cfg->max_steps) {
      bool final_step = false;
      T cap = static_cast<T>(cfg->max_dt);
      if (cfg->stop_rule == CEO_STOP_MOMENTUM_OR_TMAX) {
        const double rem = cfg->t_max - t;
        if (rem <= cfg->max_dt) {
          cap = static_cast<T>(rem);
          final_step = true;
        }
        vold_.x = vel_.x;
        vold_.y = vel_.y;
        vold_.z = vel_.z;
      }
      const T h = step(done, cap, cfg, attempts_total);
      if (h < T(0)) return -1;
      ++iter;
      if (cfg->stop_rule == CEO_STOP_PE_FRACTION) {
        t += static_cast<double>(h);
        potential(pos_, pe);
        for (int l = 0; l < K; ++l) {
          if (!done[l] && static_cast<double>(pe[l]) <= pes_[l]) {
            done[l] = true;
            out[l].status = CEO_OK_PE;
            out[l].steps = static_cast<uint32_t>(iter);
            out[l].t = t;
            latch(l);
            ++n_done;
          }
        }
      } else {
        final_step = final_step && (h >= cap || t + static_cast<double>(h) >= cfg->t_max);
        t = final_step ? cfg->t_max : t + static_cast<double>(h);
        for (int l = 0; l < K; ++l) {
          if (done[l]) continue;
          bool settled = !final_step;  // not tested on the step landing on t_max
          for (int i = 0; i < n_ && settled; ++i) {
            const int o = i * K + l;
            const T dx = vel_.x[o] - vold_.x[o], dy = vel_.y[o] - vold_.y[o],
                    dz = vel_.z[o] - vold_.z[o];
            const T num = dx * dx + dy * dy + dz * dz;
            const T den = vel_.x[o] * vel_.x[o] + vel_.y[o] * vel_.y[o] + vel_.z[o] * vel_.z[o];
            settled = num < tol2 * den;
          }
          if (settled || final_step) {
            done[l] = true;
            out[l].status = settled ? CEO_OK_MOMENTUM : CEO_OK_TMAX;
            out[l].steps = static_cast<uint32_t>(iter);
            out[l].t = t;
            latch(l);
            ++n_done;
          }
        }
      }
    }
    for (int l = 0; l < K; ++l) {
      if (!done[l]) {
        out[l].steps = static_cast<uint32_t>(iter);
        out[l].t = t;
        latch(l);
      }
    }
    // Finalize (batched_integrator.hpp:172-189) in double.
    for (int l = 0; l < K; ++l) {
      double ke = 0.0;
      for (int i = 0; i < n_; ++i) {
        const double vx = fin_.x[i * K + l], vy = fin_.y[i * K + l], vz = fin_.z[i * K + l];
        ke += 0.5 * md_[i] * (vx * vx + vy * vy + vz * vz);
      }
      double s = 1.0;
      if (cfg->redistribute && ke > 0.0) s = std::sqrt(e0_[l] / ke);
      for (int i = 0; i < n_; ++i) {
        const double m = md_[i] * s;
        out[l].mom[3 * i] = m * static_cast<double>(fin_.x[i * K + l]);
        out[l].mom[3 * i + 1] = m * static_cast<double>(fin_.y[i * K + l]);
        out[l].mom[3 * i + 2] = m * static_cast<double>(fin_.z[i * K + l]);
      }
      out[l].eerr = (ke + fin_pe_[l] - e0_[l]) / e0_[l];
    }
    return static_cast<int64_t>(iter);
  }

 private:
  // batched_force.hpp:31-84
  void accel(Soa3& src, Soa3& dst) {
    const int n = n_;
    T* __restrict ax = dst.x.data();
    T* __restrict ay = dst.y.data();
    T* __restrict az = dst.z.data();
    const T* __restrict px = src.x.data();
    const T* __restrict py = src.y.data();
    const T* __restrict pz = src.z.data();
    for (int q = 0; q < n * K; ++q) ax[q] = ay[q] = az[q] = T(0);
    for (int i = 0; i < n; ++i) {
      T axi[K], ayi[K], azi[K];
      for (int l = 0; l < K; ++l) {
        axi[l] = ax[i * K + l];
        ayi[l] = ay[i * K + l];
        azi[l] = az[i * K + l];
      }
      for (int j = i + 1; j < n; ++j) {
        const T pcij = pc_[i * n + j];
#pragma GCC ivdep
        for (int l = 0; l < K; ++l) {
          const T rx = px[i * K + l] - px[j * K + l];
          const T ry = py[i * K + l] - py[j * K + l];
          const T rz = pz[i * K + l] - pz[j * K + l];
          const T d2 = std::fma(rx, rx, std::fma(ry, ry, rz * rz));
          const T inv = T(1) / std::sqrt(d2);
          const T inv3 = inv / d2;
          const T sc = pcij * inv3;
          const T fx = rx * sc, fy = ry * sc, fz = rz * sc;
          axi[l] += fx;
          ayi[l] += fy;
          azi[l] += fz;
          ax[j * K + l] -= fx;
          ay[j * K + l] -= fy;
          az[j * K + l] -= fz;
        }
      }
      for (int l = 0; l < K; ++l) {
        ax[i * K + l] = axi[l];
        ay[i * K + l] = ayi[l];
        az[i * K + l] = azi[l];
      }
    }
    for (int i = 0; i < n; ++i) {
      const T m = im_[i];
      for (int l = 0; l < K; ++l) {
        ax[i * K + l] *= m;
        ay[i * K + l] *= m;
        az[i * K + l] *= m;
      }
    }
  }

  // batched_force.hpp:89-110
  void potential(Soa3& src, T* pe) {
    T u[K];
    for (int l = 0; l < K; ++l) u[l] = T(0);
    for (int i = 0; i < n_; ++i)
      for (int j = i + 1; j < n_; ++j) {
        const T pcij = pc_[i * n_ + j];
        for (int l = 0; l < K; ++l) {
          const T rx = src.x[i * K + l] - src.x[j * K + l];
          const T ry = src.y[i * K + l] - src.y[j * K + l];
          const T rz = src.z[i * K + l] - src.z[j * K + l];
          const T d2 = std::fma(rx, rx, std::fma(ry, ry, rz * rz));
          u[l] = std::fma(pcij, T(1) / std::sqrt(d2), u[l]);
        }
      }
    for (int l = 0; l < K; ++l) pe[l] = u[l];
  }

  void latch(int l) {
    for (int i = 0; i < n_; ++i) {
      fin_.x[i * K + l] = vel_.x[i * K + l];
      fin_.y[i * K + l] = vel_.y[i * K + l];
      fin_.z[i * K + l] = vel_.z[i * K + l];
    }
    // FP64 potential at the latch point, for the energy diagnostic only.
    double u = 0.0;
    for (int i = 0; i < n_; ++i)
      for (int j = i + 1; j < n_; ++j) {
        const double rx = double(pos_.x[i * K + l]) - double(pos_.x[j * K + l]);
        const double ry = double(pos_.y[i * K + l]) - double(pos_.y[j * K + l]);
        const double rz = double(pos_.z[i * K + l]) - double(pos_.z[j * K + l]);
        u += double(pc_[i * n_ + j]) / std::sqrt(rx * rx + ry * ry + rz * rz);
      }
    fin_pe_[l] = u;
  }

  // batched_integrator.hpp:228-343 -- one accepted shared-dt step; returns the
  // step taken, or -1 after more than max_rej rejections.
  T step(const bool* done, T max_dt, const ceo_config* cfg, uint32_t* attempts_total) {
    const int n = n_;
    const Coef<T>& C = coef<T>();
    if (!fsal_) {
      kp_[0].x = vel_.x;
      kp_[0].y = vel_.y;
      kp_[0].z = vel_.z;
      accel(pos_, kv_[0]);
      fsal_ = true;
    }
    if (h_ <= T(0))
      h_ = (cfg->h0_mode == CEO_H0_FIXED) ? static_cast<T>(cfg->h0) : initial_step(done, max_dt);

    bool rejected_once = false;
    for (int attempt = 0;; ++attempt) {
      if (attempt > max_rej_) return T(-1);
      const T h = std::min(h_, max_dt);
      if (attempts_total) ++*attempts_total;
      for (int s = 1; s < 7; ++s) {
        for (int i = 0; i < n; ++i) {
          const int o = i * K;
          for (int c = 0; c < 3; ++c) {
            T dx[K], dv[K];
            for (int l = 0; l < K; ++l) dx[l] = dv[l] = T(0);
            for (int j = 0; j < s; ++j) {
              const T aj = C.a[s][j];
              const T* __restrict kpj = kp_[j].c(c) + o;
              const T* __restrict kvj = kv_[j].c(c) + o;
              for (int l = 0; l < K; ++l) {
                dx[l] = std::fma(kpj[l], aj, dx[l]);
                dv[l] = std::fma(kvj[l], aj, dv[l]);
              }
            }
            const T* __restrict P = pos_.c(c) + o;
            const T* __restrict V = vel_.c(c) + o;
            T* __restrict ST = stage_.c(c) + o;
            T* __restrict KP = kp_[s].c(c) + o;
            for (int l = 0; l < K; ++l) {
              ST[l] = std::fma(dx[l], h, P[l]);
              KP[l] = std::fma(dv[l], h, V[l]);
            }
          }
        }
        accel(stage_, kv_[s]);
      }
      T err_sq[K];
      for (int l = 0; l < K; ++l) err_sq[l] = T(0);
      for (int i = 0; i < n; ++i) {
        const int o = i * K;
        for (int c = 0; c < 3; ++c) {
          T ep[K], ev[K];
          for (int l = 0; l < K; ++l) ep[l] = ev[l] = T(0);
          for (int s = 0; s < 7; ++s) {
            const T es = C.e[s];
            const T* __restrict kps = kp_[s].c(c) + o;
            const T* __restrict kvs = kv_[s].c(c) + o;
            for (int l = 0; l < K; ++l) {
              ep[l] = std::fma(kps[l], es, ep[l]);
              ev[l] = std::fma(kvs[l], es, ev[l]);
            }
          }
          const T* __restrict P = pos_.c(c) + o;
          const T* __restrict V = vel_.c(c) + o;
          const T* __restrict ST = stage_.c(c) + o;
          const T* __restrict K6 = kp_[6].c(c) + o;
          for (int l = 0; l < K; ++l) {
            const T scp = atol_ + rtol_ * std::max(std::abs(P[l]), std::abs(ST[l]));
            const T rp = (ep[l] * h) / scp;
            err_sq[l] = err_sq[l] + rp * rp;
            const T scv = atol_ + rtol_ * std::max(std::abs(V[l]), std::abs(K6[l]));
            const T rv = (ev[l] * h) / scv;
            err_sq[l] = err_sq[l] + rv * rv;
          }
        }
      }
      const T inv6n = T(1) / static_cast<T>(6 * n);
      // std::max(acc, NaN) keeps acc, so a NaN lane never blocks the batch
      // (batched_integrator.hpp:314-318); kept as the reference has it.
      T max_err = T(0);
      for (int l = 0; l < K; ++l) {
        const T en = std::sqrt(err_sq[l] * inv6n);
        if (!done[l]) max_err = std::max(max_err, en);
      }
      if (max_err <= T(1)) {
        pos_.swap(stage_);
        vel_.swap(kp_[6]);
        kv_[0].swap(kv_[6]);
        kp_[0].x = vel_.x;
        kp_[0].y = vel_.y;
        kp_[0].z = vel_.z;
        const double e = static_cast<double>(max_err);
        T f = (max_err == T(0)) ? static_cast<T>(kMaxFactor)
                                : static_cast<T>(std::min(kMaxFactor, kSafety * std::pow(e, kExponent)));
        if (rejected_once) f = std::min(f, T(1));
        h_ = std::min(h * f, max_dt);
        return h;
      }
      const double e = static_cast<double>(max_err);
      const T f = static_cast<T>(std::max(kMinFactor, kSafety * std::pow(e, kExponent)));
      h_ = h * f;
      rejected_once = true;
    }
  }

  // batched_integrator.hpp:348-420
  T initial_step(const bool* done, T max_dt) {
    const int n = n_;
    const T invden = T(1) / static_cast<T>(6 * n);
    T d0sq[K], d1sq[K], d2sq[K], d1[K], d2[K], h0[K];
    for (int l = 0; l < K; ++l) d0sq[l] = d1sq[l] = d2sq[l] = T(0);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        const int o = i * K;
        for (int l = 0; l < K; ++l) {
          const T x = pos_.c(c)[o + l], v = vel_.c(c)[o + l];
          const T sx = atol_ + rtol_ * std::abs(x);
          const T sv = atol_ + rtol_ * std::abs(v);
          const T rx = x / sx, rv = v / sv;
          d0sq[l] = d0sq[l] + (rx * rx + rv * rv);
          const T gp = kp_[0].c(c)[o + l] / sx;
          const T gv = kv_[0].c(c)[o + l] / sv;
          d1sq[l] = d1sq[l] + (gp * gp + gv * gv);
        }
      }
    for (int l = 0; l < K; ++l) {
      const T d0 = std::sqrt(d0sq[l] * invden);
      d1[l] = std::sqrt(d1sq[l] * invden);
      h0[l] = (d0 < T(1e-5) || d1[l] < T(1e-5)) ? T(1e-6) : T(0.01) * d0 / d1[l];
    }
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c)
        for (int l = 0; l < K; ++l)
          stage_.c(c)[i * K + l] = std::fma(h0[l], kp_[0].c(c)[i * K + l], pos_.c(c)[i * K + l]);
    accel(stage_, probe_);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        const int o = i * K;
        for (int l = 0; l < K; ++l) {
          const T sx = atol_ + rtol_ * std::abs(pos_.c(c)[o + l]);
          const T sv = atol_ + rtol_ * std::abs(vel_.c(c)[o + l]);
          const T dvp = kv_[0].c(c)[o + l] * h0[l];
          const T dap = probe_.c(c)[o + l] - kv_[0].c(c)[o + l];
          const T rp = dvp / sx, ra = dap / sv;
          d2sq[l] = d2sq[l] + (rp * rp + ra * ra);
        }
      }
    for (int l = 0; l < K; ++l) d2[l] = std::sqrt(d2sq[l] * invden) / h0[l];
    T hmin = max_dt;
    for (int l = 0; l < K; ++l) {
      const double d1l = static_cast<double>(d1[l]);
      const double d2l = static_cast<double>(d2[l]);
      const T h1 = (d1l <= 1e-15 && d2l <= 1e-15)
                       ? std::max(T(1e-6), h0[l] * T(1e-3))
                       : static_cast<T>(std::pow(0.01 / std::max(d1l, d2l), 1.0 / 5.0));
      const T hl = std::min({T(100) * h0[l], h1, max_dt});
      if (!done[l]) hmin = std::min(hmin, hl);
    }
    return hmin;
  }

  int n_;
  T rtol_, atol_;
  int max_rej_;
  T h_{0};
  bool fsal_{false};
  std::vector<T> pc_, im_;
  std::vector<double> md_;
  Soa3 pos_, vel_, stage_, probe_, fin_, vold_;
  Soa3 kp_[7], kv_[7];
  double pe0_[K], pes_[K], e0_[K], fin_pe_[K];
};

}  // namespace ceo
