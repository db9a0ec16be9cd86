// CPU oracle, lockstep batched engine entry points -- TEST INFRASTRUCTURE ONLY.
// See batched.hpp for what is restated (include/coulomb/batched_integrator.hpp).
#include <cstring>
#include <vector>

#include "batched.hpp"

namespace ceo {
namespace {

template <class T, int K>
int batched_run(const ceo_molecule* mol, const ceo_config* cfg, const double* pos, size_t n_sims,
                ceo_results* out, uint64_t* batch_steps) {
  if (mol->n < 1 || mol->n > 16) return -1;
  const int n = mol->n;
  Lockstep<T, K> eng(mol, static_cast<T>(cfg->rtol), static_cast<T>(cfg->atol),
                     cfg->max_rejections);
  std::vector<typename Lockstep<T, K>::LaneOut> lo(K);
  uint64_t total = 0;
  for (size_t b = 0; b < n_sims; b += K) {
    for (int l = 0; l < K; ++l) {
      const size_t sim = (b + l < n_sims) ? b + l : n_sims - 1;
      eng.set_lane(l, pos + sim * n * 3);
    }
    uint32_t att = 0;
    const int64_t it = eng.run(cfg, lo.data(), &att);
    if (it < 0) return -2;  // rejection storm: the reference throws runtime_error
    total += static_cast<uint64_t>(it);
    for (int l = 0; l < K && b + l < n_sims; ++l) {
      const size_t sim = b + l;
      std::memcpy(out->momenta + sim * n * 3, lo[l].mom, sizeof(double) * 3 * n);
      out->steps[sim] = lo[l].steps;
      out->status[sim] = lo[l].status;
      if (out->attempts) out->attempts[sim] = att;
      if (out->t_final) out->t_final[sim] = lo[l].t;
      if (out->energy_err) out->energy_err[sim] = lo[l].eerr;
    }
  }
  if (batch_steps) *batch_steps = total;
  return 0;
}

template <class T>
int dispatch(const ceo_molecule* mol, const ceo_config* cfg, const double* pos, size_t n_sims,
             int32_t lanes, ceo_results* out, uint64_t* bs) {
  switch (lanes) {
    case 1: return batched_run<T, 1>(mol, cfg, pos, n_sims, out, bs);
    case 4: return batched_run<T, 4>(mol, cfg, pos, n_sims, out, bs);
    case 8: return batched_run<T, 8>(mol, cfg, pos, n_sims, out, bs);
    case 16: return batched_run<T, 16>(mol, cfg, pos, n_sims, out, bs);
    default: return -1;  // the reference throws invalid_argument for a bad width
  }
}

}  // namespace
}  // namespace ceo

extern "C" {

int ceo_batched_run_f32(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                        size_t n_sims, int32_t lanes, ceo_results* out, uint64_t* batch_steps) {
  return ceo::dispatch<float>(mol, cfg, pos, n_sims, lanes, out, batch_steps);
}

int ceo_batched_run_f64(const ceo_molecule* mol, const ceo_config* cfg, const double* pos,
                        size_t n_sims, int32_t lanes, ceo_results* out, uint64_t* batch_steps) {
  return ceo::dispatch<double>(mol, cfg, pos, n_sims, lanes, out, batch_steps);
}

}  // extern "C"

// batched::accelerations<float> / potential_energy<float> (batched_force.hpp:31-110)
// on the reference's SoA layout p[i*k + lane], any k: one lane at a time with the
// reference's per-lane arithmetic (MulAdd -> fma, Div(1, Sqrt), Div by d2).
extern "C" void ceo_soa_accelerations_f32(size_t n, size_t k, const float* px, const float* py,
                                          const float* pz, const float* pc, const float* im,
                                          float* ax, float* ay, float* az) {
  for (size_t q = 0; q < n * k; ++q) ax[q] = ay[q] = az[q] = 0.0f;
  for (size_t l = 0; l < k; ++l) {
    for (size_t i = 0; i < n; ++i) {
      float axi = ax[i * k + l], ayi = ay[i * k + l], azi = az[i * k + l];
      for (size_t j = i + 1; j < n; ++j) {
        const float rx = px[i * k + l] - px[j * k + l];
        const float ry = py[i * k + l] - py[j * k + l];
        const float rz = pz[i * k + l] - pz[j * k + l];
        const float d2 = std::fma(rx, rx, std::fma(ry, ry, rz * rz));
        const float inv = 1.0f / std::sqrt(d2);
        const float inv3 = inv / d2;
        const float s = pc[i * n + j] * inv3;
        const float fx = rx * s, fy = ry * s, fz = rz * s;
        axi += fx;
        ayi += fy;
        azi += fz;
        ax[j * k + l] -= fx;
        ay[j * k + l] -= fy;
        az[j * k + l] -= fz;
      }
      ax[i * k + l] = axi;
      ay[i * k + l] = ayi;
      az[i * k + l] = azi;
    }
    for (size_t i = 0; i < n; ++i) {
      ax[i * k + l] *= im[i];
      ay[i * k + l] *= im[i];
      az[i * k + l] *= im[i];
    }
  }
}

extern "C" void ceo_soa_potential_energy_f32(size_t n, size_t k, const float* px, const float* py,
                                             const float* pz, const float* pc, float* pe) {
  for (size_t l = 0; l < k; ++l) {
    float u = 0.0f;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = i + 1; j < n; ++j) {
        const float rx = px[i * k + l] - px[j * k + l];
        const float ry = py[i * k + l] - py[j * k + l];
        const float rz = pz[i * k + l] - pz[j * k + l];
        const float d2 = std::fma(rx, rx, std::fma(ry, ry, rz * rz));
        u = std::fma(pc[i * n + j], 1.0f / std::sqrt(d2), u);
      }
    pe[l] = u;
  }
}
