// CPU baseline -- MEASUREMENT INFRASTRUCTURE ONLY (see ce_oracle.h).
//
// The reference's production engine is BatchedRK45<float> at the host SIMD width
// (examples/generate_dataset.cpp:163-196) on one thread.  Its Highway build
// cannot be made here, so this file restates it with explicit SIMD intrinsics:
// K = 16 lanes on AVX-512, 8 on AVX2+FMA, 4 on SSE2, one simulation per lane,
// SoA over lanes; per lane the arithmetic of batched_force.hpp:31-110 and
// batched_integrator.hpp:228-433 (MulAdd -> FMA, IEEE Sqrt and Div, the
// controller's pow in double on the worst active lane).  It is compiled once per
// ISA (CEO_ISA_TAG set by the Makefile) and run as one independent instance per
// thread over disjoint sim ranges, timed like
// bench/bench_batched_integrator.cpp:139-175 (pre-sampled geometries, median of
// the timed passes).  tests/ check its results against batched.hpp.
#include <immintrin.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "ce_oracle.h"
#include "oracle_common.hpp"

#ifndef CEO_ISA_TAG
#error "CEO_ISA_TAG must be defined by the build"
#endif
#define CEO_CAT2(a, b) a##b
#define CEO_CAT(a, b) CEO_CAT2(a, b)

namespace {

#if defined(__AVX512F__)
constexpr int K = 16;
using V = __m512;
inline V vset(float a) { return _mm512_set1_ps(a); }
inline V vzero() { return _mm512_setzero_ps(); }
inline V vadd(V a, V b) { return _mm512_add_ps(a, b); }
inline V vsub(V a, V b) { return _mm512_sub_ps(a, b); }
inline V vmul(V a, V b) { return _mm512_mul_ps(a, b); }
inline V vdiv(V a, V b) { return _mm512_div_ps(a, b); }
inline V vfma(V a, V b, V c) { return _mm512_fmadd_ps(a, b, c); }
inline V vsqrt(V a) { return _mm512_sqrt_ps(a); }
inline V vmax(V a, V b) { return _mm512_max_ps(a, b); }
inline V vabs(V a) { return _mm512_abs_ps(a); }
inline void vstore(float* p, V a) { _mm512_storeu_ps(p, a); }
#elif defined(__AVX2__) && defined(__FMA__)
constexpr int K = 8;
using V = __m256;
inline V vset(float a) { return _mm256_set1_ps(a); }
inline V vzero() { return _mm256_setzero_ps(); }
inline V vadd(V a, V b) { return _mm256_add_ps(a, b); }
inline V vsub(V a, V b) { return _mm256_sub_ps(a, b); }
inline V vmul(V a, V b) { return _mm256_mul_ps(a, b); }
inline V vdiv(V a, V b) { return _mm256_div_ps(a, b); }
inline V vfma(V a, V b, V c) { return _mm256_fmadd_ps(a, b, c); }
inline V vsqrt(V a) { return _mm256_sqrt_ps(a); }
inline V vmax(V a, V b) { return _mm256_max_ps(a, b); }
inline V vabs(V a) { return _mm256_andnot_ps(_mm256_set1_ps(-0.0f), a); }
inline void vstore(float* p, V a) { _mm256_storeu_ps(p, a); }
#else
constexpr int K = 4;
using V = __m128;
inline V vset(float a) { return _mm_set1_ps(a); }
inline V vzero() { return _mm_setzero_ps(); }
inline V vadd(V a, V b) { return _mm_add_ps(a, b); }
inline V vsub(V a, V b) { return _mm_sub_ps(a, b); }
inline V vmul(V a, V b) { return _mm_mul_ps(a, b); }
inline V vdiv(V a, V b) { return _mm_div_ps(a, b); }
// No FMA unit on the SSE2 baseline: MulAdd is a multiply then an add there.
inline V vfma(V a, V b, V c) { return _mm_add_ps(_mm_mul_ps(a, b), c); }
inline V vsqrt(V a) { return _mm_sqrt_ps(a); }
inline V vmax(V a, V b) { return _mm_max_ps(a, b); }
inline V vabs(V a) { return _mm_andnot_ps(_mm_set1_ps(-0.0f), a); }
inline void vstore(float* p, V a) { _mm_storeu_ps(p, a); }
#endif

constexpr int kMaxN = 16;

// One lockstep batch of K lanes of an n-atom molecule.
class Batch {
 public:
  Batch(const ceo_molecule* m, float rtol, float atol, int max_rej)
      : n_(m->n), rtol_(rtol), atol_(atol), max_rej_(max_rej) {
    for (int i = 0; i < n_; ++i) {
      md_[i] = m->masses[i];
      im_[i] = static_cast<float>(1.0 / m->masses[i]);
      for (int j = 0; j < n_; ++j)
        pc_[i][j] = static_cast<float>(m->coulomb_k * m->charges[i] * m->charges[j]);
    }
    const ceo::Coef<float>& C = ceo::coef<float>();
    for (int s = 0; s < 7; ++s) {
      for (int j = 0; j < 7; ++j) a_[s][j] = vset(C.a[s][j]);
      e_[s] = vset(C.e[s]);
    }
  }

  void set_lane(int lane, const double* p) {
    for (int i = 0; i < n_; ++i)
      for (int c = 0; c < 3; ++c) {
        float* px = reinterpret_cast<float*>(&x_[c][i]);
        float* pv = reinterpret_cast<float*>(&v_[c][i]);
        px[lane] = static_cast<float>(p[3 * i + c]);
        pv[lane] = 0.0f;
      }
  }

  // Runs the batch to its stop; fills per-lane momenta. Returns batch iterations.
  int64_t run(const ceo_config* cfg, double* mom /* [K][n][3] */) {
    const bool pe_rule = cfg->stop_rule == CEO_STOP_PE_FRACTION;
    double pe0[K], pes[K];
    float pe[K];
    potential(x_, pe);
    int n_done = 0;
    for (int l = 0; l < K; ++l) {
      pe0[l] = pe[l];
      pes[l] = cfg->pe_stop_fraction * pe0[l];
      done_[l] = false;
      if (pe_rule && pe0[l] <= pes[l]) {
        done_[l] = true;
        latch(l);
        ++n_done;
      }
    }
    h_ = 0.0f;
    fsal_ = false;
    uint64_t iter = 0;
    double t = 0.0;
    const float tol2 = static_cast<float>(cfg->momentum_tol * cfg->momentum_tol);
    while (n_done < K && iter < cfg->max_steps) {
      float cap = static_cast<float>(cfg->max_dt);
      bool final_step = false;
      if (!pe_rule) {
        const double rem = cfg->t_max - t;
        if (rem <= cfg->max_dt) {
          cap = static_cast<float>(rem);
          final_step = true;
        }
        std::memcpy(vold_, v_, sizeof(v_));
      }
      const float h = step(cap, cfg);
      if (h < 0.0f) return -1;
      ++iter;
      if (pe_rule) {
        potential(x_, pe);
        for (int l = 0; l < K; ++l)
          if (!done_[l] && static_cast<double>(pe[l]) <= pes[l]) {
            done_[l] = true;
            latch(l);
            ++n_done;
          }
      } else {
        final_step = final_step && (h >= cap || t + static_cast<double>(h) >= cfg->t_max);
        t = final_step ? cfg->t_max : t + static_cast<double>(h);
        for (int l = 0; l < K; ++l) {
          if (done_[l]) continue;
          bool settled = !final_step;  // not tested on the step landing on t_max
          for (int i = 0; i < n_ && settled; ++i) {
            float num = 0.0f, den = 0.0f;
            for (int c = 0; c < 3; ++c) {
              const float vn = lane_of(v_[c][i], l), vo = lane_of(vold_[c][i], l);
              num += (vn - vo) * (vn - vo);
              den += vn * vn;
            }
            settled = num < tol2 * den;
          }
          if (settled || final_step) {
            done_[l] = true;
            latch(l);
            ++n_done;
          }
        }
      }
    }
    for (int l = 0; l < K; ++l)
      if (!done_[l]) latch(l);
    for (int l = 0; l < K; ++l) {
      double ke = 0.0;
      for (int i = 0; i < n_; ++i) {
        double s2 = 0.0;
        for (int c = 0; c < 3; ++c) {
          const double w = fin_[l][3 * i + c];
          s2 += w * w;
        }
        ke += 0.5 * md_[i] * s2;
      }
      const double s = (cfg->redistribute && ke > 0.0) ? std::sqrt(pe0[l] / ke) : 1.0;
      for (int i = 0; i < n_; ++i)
        for (int c = 0; c < 3; ++c)
          mom[(l * n_ + i) * 3 + c] = md_[i] * s * static_cast<double>(fin_[l][3 * i + c]);
    }
    return static_cast<int64_t>(iter);
  }

 private:
  static float lane_of(const V& v, int l) { return reinterpret_cast<const float*>(&v)[l]; }

  void latch(int l) {
    for (int i = 0; i < n_; ++i)
      for (int c = 0; c < 3; ++c) fin_[l][3 * i + c] = lane_of(v_[c][i], l);
  }

  // batched_force.hpp:31-84
  void accel(const V (&x)[3][kMaxN], V (&a)[3][kMaxN]) const {
    const int n = n_;
    for (int i = 0; i < n; ++i) a[0][i] = a[1][i] = a[2][i] = vzero();
    const V one = vset(1.0f);
    for (int i = 0; i < n; ++i) {
      V ax = a[0][i], ay = a[1][i], az = a[2][i];
      for (int j = i + 1; j < n; ++j) {
        const V rx = vsub(x[0][i], x[0][j]);
        const V ry = vsub(x[1][i], x[1][j]);
        const V rz = vsub(x[2][i], x[2][j]);
        const V d2 = vfma(rx, rx, vfma(ry, ry, vmul(rz, rz)));
        const V inv = vdiv(one, vsqrt(d2));
        const V inv3 = vdiv(inv, d2);
        const V s = vmul(vset(pc_[i][j]), inv3);
        const V fx = vmul(rx, s), fy = vmul(ry, s), fz = vmul(rz, s);
        ax = vadd(ax, fx);
        ay = vadd(ay, fy);
        az = vadd(az, fz);
        a[0][j] = vsub(a[0][j], fx);
        a[1][j] = vsub(a[1][j], fy);
        a[2][j] = vsub(a[2][j], fz);
      }
      a[0][i] = ax;
      a[1][i] = ay;
      a[2][i] = az;
    }
    for (int i = 0; i < n; ++i) {
      const V m = vset(im_[i]);
      a[0][i] = vmul(a[0][i], m);
      a[1][i] = vmul(a[1][i], m);
      a[2][i] = vmul(a[2][i], m);
    }
  }

  // batched_force.hpp:89-110
  void potential(const V (&x)[3][kMaxN], float* pe) const {
    V u = vzero();
    const V one = vset(1.0f);
    for (int i = 0; i < n_; ++i)
      for (int j = i + 1; j < n_; ++j) {
        const V rx = vsub(x[0][i], x[0][j]);
        const V ry = vsub(x[1][i], x[1][j]);
        const V rz = vsub(x[2][i], x[2][j]);
        const V d2 = vfma(rx, rx, vfma(ry, ry, vmul(rz, rz)));
        u = vfma(vset(pc_[i][j]), vdiv(one, vsqrt(d2)), u);
      }
    vstore(pe, u);
  }

  // batched_integrator.hpp:228-343, fixed-h0 or a carried step size.
  float step(float max_dt, const ceo_config* cfg) {
    const int n = n_;
    if (!fsal_) {
      std::memcpy(kp_[0], v_, sizeof(v_));
      accel(x_, kv_[0]);
      fsal_ = true;
    }
    if (h_ <= 0.0f)
      h_ = (cfg->h0_mode == CEO_H0_FIXED) ? static_cast<float>(cfg->h0) : initial_step(max_dt);
    const V rtol = vset(rtol_), atol = vset(atol_);
    const V inv6n = vset(1.0f / static_cast<float>(6 * n));
    bool rejected_once = false;
    for (int attempt = 0;; ++attempt) {
      if (attempt > max_rej_) return -1.0f;
      const float h = std::min(h_, max_dt);
      const V hv = vset(h);
      for (int s = 1; s < 7; ++s) {
        for (int i = 0; i < n; ++i)
          for (int c = 0; c < 3; ++c) {
            V dx = vzero(), dv = vzero();
            for (int j = 0; j < s; ++j) {
              dx = vfma(kp_[j][c][i], a_[s][j], dx);
              dv = vfma(kv_[j][c][i], a_[s][j], dv);
            }
            st_[c][i] = vfma(dx, hv, x_[c][i]);
            kp_[s][c][i] = vfma(dv, hv, v_[c][i]);
          }
        accel(st_, kv_[s]);
      }
      V err = vzero();
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
          V ep = vzero(), ev = vzero();
          for (int s = 0; s < 7; ++s) {
            ep = vfma(kp_[s][c][i], e_[s], ep);
            ev = vfma(kv_[s][c][i], e_[s], ev);
          }
          const V scp = vadd(atol, vmul(rtol, vmax(vabs(x_[c][i]), vabs(st_[c][i]))));
          const V rp = vdiv(vmul(ep, hv), scp);
          err = vadd(err, vmul(rp, rp));
          const V scv = vadd(atol, vmul(rtol, vmax(vabs(v_[c][i]), vabs(kp_[6][c][i]))));
          const V rv = vdiv(vmul(ev, hv), scv);
          err = vadd(err, vmul(rv, rv));
        }
      float en[K];
      vstore(en, vsqrt(vmul(err, inv6n)));
      float max_err = 0.0f;
      for (int l = 0; l < K; ++l)
        if (!done_[l]) max_err = std::max(max_err, en[l]);
      if (max_err <= 1.0f) {
        std::memcpy(x_, st_, sizeof(x_));
        std::memcpy(v_, kp_[6], sizeof(v_));
        std::memcpy(kv_[0], kv_[6], sizeof(kv_[0]));
        std::memcpy(kp_[0], v_, sizeof(v_));
        const double e = max_err;
        float f = (max_err == 0.0f)
                      ? static_cast<float>(ceo::kMaxFactor)
                      : static_cast<float>(std::min(ceo::kMaxFactor,
                                                    ceo::kSafety * std::pow(e, ceo::kExponent)));
        if (rejected_once) f = std::min(f, 1.0f);
        h_ = std::min(h * f, max_dt);
        return h;
      }
      const double e = max_err;
      h_ = h * static_cast<float>(std::max(ceo::kMinFactor, ceo::kSafety * std::pow(e, ceo::kExponent)));
      rejected_once = true;
    }
  }

  // batched_integrator.hpp:348-420 (Hairer start, min over active lanes).
  float initial_step(float max_dt) {
    const int n = n_;
    const V rtol = vset(rtol_), atol = vset(atol_);
    const V invden = vset(1.0f / static_cast<float>(6 * n));
    V d0sq = vzero(), d1sq = vzero();
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        const V sx = vadd(atol, vmul(rtol, vabs(x_[c][i])));
        const V sv = vadd(atol, vmul(rtol, vabs(v_[c][i])));
        const V rx = vdiv(x_[c][i], sx), rv = vdiv(v_[c][i], sv);
        d0sq = vadd(d0sq, vadd(vmul(rx, rx), vmul(rv, rv)));
        const V gp = vdiv(kp_[0][c][i], sx), gv = vdiv(kv_[0][c][i], sv);
        d1sq = vadd(d1sq, vadd(vmul(gp, gp), vmul(gv, gv)));
      }
    float d0[K], d1[K], h0[K], d2[K];
    vstore(d0, vsqrt(vmul(d0sq, invden)));
    vstore(d1, vsqrt(vmul(d1sq, invden)));
    for (int l = 0; l < K; ++l)
      h0[l] = (d0[l] < 1e-5f || d1[l] < 1e-5f) ? 1e-6f : 0.01f * d0[l] / d1[l];
    V h0v;
    std::memcpy(&h0v, h0, sizeof(h0v));
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) st_[c][i] = vfma(h0v, kp_[0][c][i], x_[c][i]);
    accel(st_, probe_);
    V d2sq = vzero();
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        const V sx = vadd(atol, vmul(rtol, vabs(x_[c][i])));
        const V sv = vadd(atol, vmul(rtol, vabs(v_[c][i])));
        const V rp = vdiv(vmul(kv_[0][c][i], h0v), sx);
        const V ra = vdiv(vsub(probe_[c][i], kv_[0][c][i]), sv);
        d2sq = vadd(d2sq, vadd(vmul(rp, rp), vmul(ra, ra)));
      }
    vstore(d2, vdiv(vsqrt(vmul(d2sq, invden)), h0v));
    float hmin = max_dt;
    for (int l = 0; l < K; ++l) {
      const double d1l = d1[l], d2l = d2[l];
      const float h1 = (d1l <= 1e-15 && d2l <= 1e-15)
                           ? std::max(1e-6f, h0[l] * 1e-3f)
                           : static_cast<float>(std::pow(0.01 / std::max(d1l, d2l), 0.2));
      const float hl = std::min({100.0f * h0[l], h1, max_dt});
      if (!done_[l]) hmin = std::min(hmin, hl);
    }
    return hmin;
  }

  int n_;
  float rtol_, atol_;
  int max_rej_;
  float h_ = 0.0f;
  bool fsal_ = false;
  bool done_[K];
  double md_[kMaxN];
  float im_[kMaxN];
  float pc_[kMaxN][kMaxN];
  V a_[7][7], e_[7];
  V x_[3][kMaxN], v_[3][kMaxN], st_[3][kMaxN], vold_[3][kMaxN];
  V kp_[7][3][kMaxN], kv_[7][3][kMaxN], probe_[3][kMaxN];
  float fin_[K][3 * kMaxN];
};

}  // namespace

extern "C" int CEO_CAT(ceo_cpu_baseline_, CEO_ISA_TAG)(const ceo_molecule* mol,
                                                       const ceo_config* cfg, const double* pos,
                                                       size_t n_sims, int32_t n_threads,
                                                       int32_t reps, double* median_seconds,
                                                       uint64_t* total_batch_steps,
                                                       double* checksum) {
  if (n_threads < 1 || reps < 1 || mol->n < 1 || mol->n > kMaxN) return 0;
  const int n = mol->n;
  const size_t n_batches = n_sims / K;  // whole batches only
  std::vector<uint64_t> steps(n_threads, 0);
  std::vector<double> sums(n_threads, 0.0);
  auto worker = [&](int tid) {
    Batch* b = new Batch(mol, static_cast<float>(cfg->rtol), static_cast<float>(cfg->atol),
                         cfg->max_rejections);
    std::vector<double> mom(static_cast<size_t>(K) * n * 3);
    const size_t b0 = n_batches * tid / n_threads, b1 = n_batches * (tid + 1) / n_threads;
    uint64_t st = 0;
    double cs = 0.0;
    for (size_t q = b0; q < b1; ++q) {
      for (int l = 0; l < K; ++l) b->set_lane(l, pos + (q * K + l) * n * 3);
      const int64_t it = b->run(cfg, mom.data());
      st += static_cast<uint64_t>(it < 0 ? 0 : it);
      cs += mom[0];
    }
    steps[tid] = st;
    sums[tid] = cs;
    delete b;
  };
  std::vector<double> secs;
  for (int r = 0; r < reps; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t) th.emplace_back(worker, t);
    for (auto& t : th) t.join();
    secs.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  std::sort(secs.begin(), secs.end());
  const size_t m = secs.size();
  *median_seconds = (m % 2) ? secs[m / 2] : 0.5 * (secs[m / 2 - 1] + secs[m / 2]);
  uint64_t tot = 0;
  double cs = 0.0;
  for (int t = 0; t < n_threads; ++t) {
    tot += steps[t];
    cs += sums[t];
  }
  if (total_batch_steps) *total_batch_steps = tot;
  if (checksum) *checksum = cs;
  return K;
}

// Results of the baseline for checking against batched.hpp: runs the sims in
// whole batches (n_sims a multiple of the lane count) and writes momenta.
extern "C" int CEO_CAT(ceo_cpu_baseline_momenta_, CEO_ISA_TAG)(const ceo_molecule* mol,
                                                               const ceo_config* cfg,
                                                               const double* pos, size_t n_sims,
                                                               double* mom_out) {
  if (mol->n < 1 || mol->n > kMaxN || n_sims % K != 0) return 0;
  Batch* b = new Batch(mol, static_cast<float>(cfg->rtol), static_cast<float>(cfg->atol),
                       cfg->max_rejections);
  const int n = mol->n;
  for (size_t q = 0; q < n_sims / K; ++q) {
    for (int l = 0; l < K; ++l) b->set_lane(l, pos + (q * K + l) * n * 3);
    b->run(cfg, mom_out + q * K * n * 3);
  }
  delete b;
  return K;
}
