// CPU oracle shared pieces -- TEST INFRASTRUCTURE ONLY (see ce_oracle.h).
//
// The Dormand-Prince 5(4) tableau, error weights and controller constants as
// printed in the reference (src/integrators.cpp:203-226, identical in
// include/coulomb/batched_integrator.hpp:196-211), held in double and narrowed to
// T per use exactly as `static constexpr Real k*[] = {p/q, ...}` narrows them.
#pragma once

#include <cmath>

namespace ceo {

struct DP5Table {
  double a[7][7];  // a[s][j], stages s = 1..6 (row 6 = 5th-order weights b, FSAL)
  double e[7];     // b - b* (5th minus 4th order)
};

inline const DP5Table& dp5_table() {
  static const DP5Table t = [] {
    DP5Table r{};
    r.a[1][0] = 1.0 / 5;
    r.a[2][0] = 3.0 / 40;
    r.a[2][1] = 9.0 / 40;
    r.a[3][0] = 44.0 / 45;
    r.a[3][1] = -56.0 / 15;
    r.a[3][2] = 32.0 / 9;
    r.a[4][0] = 19372.0 / 6561;
    r.a[4][1] = -25360.0 / 2187;
    r.a[4][2] = 64448.0 / 6561;
    r.a[4][3] = -212.0 / 729;
    r.a[5][0] = 9017.0 / 3168;
    r.a[5][1] = -355.0 / 33;
    r.a[5][2] = 46732.0 / 5247;
    r.a[5][3] = 49.0 / 176;
    r.a[5][4] = -5103.0 / 18656;
    r.a[6][0] = 35.0 / 384;
    r.a[6][1] = 0.0;
    r.a[6][2] = 500.0 / 1113;
    r.a[6][3] = 125.0 / 192;
    r.a[6][4] = -2187.0 / 6784;
    r.a[6][5] = 11.0 / 84;
    const double e[7] = {71.0 / 57600,      0.0,        -71.0 / 16695, 71.0 / 1920,
                         -17253.0 / 339200, 22.0 / 525, -1.0 / 40};
    for (int s = 0; s < 7; ++s) r.e[s] = e[s];
    return r;
  }();
  return t;
}

template <class T>
struct Coef {
  T a[7][7];
  T e[7];
};

template <class T>
inline const Coef<T>& coef() {
  static const Coef<T> c = [] {
    Coef<T> r{};
    const DP5Table& t = dp5_table();
    for (int s = 0; s < 7; ++s) {
      for (int j = 0; j < 7; ++j) r.a[s][j] = static_cast<T>(t.a[s][j]);
      r.e[s] = static_cast<T>(t.e[s]);
    }
    return r;
  }();
  return c;
}

// Controller constants (src/integrators.cpp:220-226).
constexpr double kSafety = 0.9;
constexpr double kMinFactor = 0.2;
constexpr double kMaxFactor = 10.0;
constexpr double kExponent = -1.0 / 5.0;

// Momentum-change stop test: every atom's |v_new - v_old|^2 < tol^2 |v_new|^2.
template <class T, class Vec>
inline bool momentum_settled(int n, const Vec& vold, const Vec& vnew, T tol2) {
  for (int i = 0; i < n; ++i) {
    const T dx = vnew[i].x - vold[i].x, dy = vnew[i].y - vold[i].y, dz = vnew[i].z - vold[i].z;
    const T num = dx * dx + dy * dy + dz * dz;
    const T den = vnew[i].x * vnew[i].x + vnew[i].y * vnew[i].y + vnew[i].z * vnew[i].z;
    if (!(num < tol2 * den)) return false;
  }
  return true;
}

}  // namespace ceo
